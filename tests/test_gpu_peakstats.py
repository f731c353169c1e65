"""GPU peak statistics against the oracle (gtrap/detect_peaks.py and the
statistics block of gtrap/gtls.py:244-275 restated): peak indices, counts
and the top peaks exactly; statistics to 1e-12 relative."""
import numpy as np
import pytest

from oracle import peakstats as op

pytestmark = pytest.mark.gpu


def _series(nt, nq, seed, ties=False, nans=False):
    rng = np.random.default_rng(seed)
    x = (np.abs(rng.standard_normal((nt, nq))) * (rng.random((nt, nq)) < 0.9)).astype(np.float32)
    x += np.sin(np.arange(nt) / 37.0)[:, None].astype(np.float32)
    if ties:
        x = np.round(x, 1).astype(np.float32)
    if nans:
        x[rng.random((nt, nq)) < 0.001] = np.nan
    t0 = (np.arange(nt) * 0.0204).astype(np.float32)[:, None].repeat(nq, 1)
    return x, t0


@pytest.mark.parametrize("nt,nq,mpd,ties", [(300, 3, 10, False), (4096, 5, 48, False),
                                            (20000, 2, 48, True), (72000, 2, 10, False),
                                            (5000, 300, 48, False)])
def test_peakstats_matches_oracle(gpu, nt, nq, mpd, ties):
    import torch
    from gtrap_amd import peakstats
    x, t0 = _series(nt, nq, seed=nt + nq, ties=ties)
    out = peakstats.get_peakstats(torch.from_numpy(x).cuda(), nt, nq, mpd=mpd,
                                  tlst0=torch.from_numpy(t0).cuda())
    for q in range(nq):
        w = op.peakstats(x[:, q], t0[:, q], mpd=mpd)
        assert out["npeak"][q] == w["npeak"] and out["lent"][q] == w["lent"]
        np.testing.assert_array_equal(out["top"][q][:len(w["top"])], w["top"])
        for k in ("maxsn", "median", "std", "far", "diff", "pssn", "pinterval"):
            np.testing.assert_allclose(out[k][q], w[k], rtol=1e-12, err_msg=k)


def test_peakstats_nan_series(gpu):
    import torch
    from gtrap_amd import peakstats
    x, t0 = _series(3000, 4, seed=5, nans=True)
    out = peakstats.get_peakstats(torch.from_numpy(x).cuda(), 3000, 4, mpd=10)
    for q in range(4):
        w = op.peakstats(x[:, q], None, mpd=10)
        assert out["npeak"][q] == w["npeak"]
        np.testing.assert_array_equal(out["top"][q], w["top"])
        if np.isnan(x[:, q]).any():
            assert np.isnan(out["maxsn"][q]) and np.isnan(out["median"][q])
        np.testing.assert_allclose(out["std"][q], w["std"], rtol=1e-12)
