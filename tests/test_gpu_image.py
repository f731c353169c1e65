"""GPU image assembly against the oracle (gtrap/read_keplerlc.py:23-57,123-146
restated in oracle/image.py): bit for bit."""
import numpy as np
import pytest

from oracle import image as oi

pytestmark = pytest.mark.gpu


def _curves(nq, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(nq):
        m = int(rng.integers(n // 3, n - 5))
        c0 = int(rng.integers(1000, 50000))
        cno = c0 + np.sort(rng.choice(n - 2, m, replace=False)).astype(np.int64)
        cno[0] = c0
        t = 120.5 + (cno - c0) * 0.0204 + 1e-5 * rng.standard_normal(m)
        f = 1.0 + 1e-3 * rng.standard_normal(m)
        f[rng.random(m) < 0.02] = np.nan
        out.append((t, f, cno))
    return out


@pytest.mark.parametrize("nq,n,offt", [(1, 300, "t[0]"), (7, 2048, "t[0]"), (70, 5000, "0"),
                                       (300, 1024, "t[0]")])
def test_image_matches_oracle(gpu, nq, n, offt):
    from gtrap_amd import image
    curves = _curves(nq, n, seed=nq)
    lc, tu, n_, ntrue, nq_, inval, t0 = image.load_curves(curves, n=n, offt=offt)
    wl, wt, wn, w0 = oi.build(curves, n, offt=offt)
    np.testing.assert_array_equal(lc.cpu().numpy(), wl)
    np.testing.assert_array_equal(tu.cpu().numpy(), wt)
    np.testing.assert_array_equal(ntrue.cpu().numpy().view(np.uint32), wn)
    np.testing.assert_array_equal(t0, w0)
    assert n_ == n and nq_ == nq and inval == -1000.0


def test_image_out_of_range_raises(gpu):
    from gtrap_amd import image
    t = np.array([0.0, 1.0])
    with pytest.raises(ValueError):
        image.load_curves([(t, np.ones(2), np.array([5, 900]))], n=100)


def test_image_feeds_geebls_offset(gpu):
    """The built image goes straight into get_offset (gfilter.py:94-101)."""
    from gtrap_amd import gfilter, image
    curves = _curves(5, 1024, seed=3)
    lc, tu, n, ntrue, nq, inval, t0 = image.load_curves(curves, n=1024)
    off, _ = gfilter.get_offset(n, nq, lc, tu.cpu().numpy())
    wl, wt, _, _ = oi.build(curves, 1024)
    want = np.where(wt >= 0, wl, 0).astype(np.float64).mean(0)
    np.testing.assert_allclose(np.asarray(off, np.float64), want, rtol=1e-6)
