"""Pins the gtls oracle (include/gtls_simple.h:1-406, include/gtls_absolute.h)."""
import numpy as np

from oracle import gtls as og
from oracle import tlsfold as ot


def _series(n=200, seed=0, dip=None):
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64) * 1.0
    x = 1e-2 * rng.standard_normal(n)
    if dip:
        k0, w, h = dip
        x[np.abs(t - k0) < w] += h
    return t.astype(np.float32), x.astype(np.float32)


def test_matches_tls_py_statistic_when_scoop_is_inside_window():
    """nsc < W/deltat: no sample precedes -W/2 (nSx = 0, the A.1 quirk is void),
    every scoop sample lies in [-W/2, W/2]; the float32 restatement then equals
    tls.py's get_chi2Hmax (float64) with the reference's sign (S/N = +H/...)."""
    t, x = _series(dip=(100, 2, 0.05))
    lc, tu = x[:, None], t[:, None]
    nsc, nl, wmax = 9, 5, 9.0  # window [-4.5, 4.5] holds exactly the 9 scoop samples
    out = og.gtls(lc, tu, 1, nl, nsc, wmax, 0.0, 1.0)
    for k in (50, 100, 150):
        best = 0.0
        for m in range(1, nl):
            sn, H, n = ot.trial_numpy(t, x, 1e9, k, wmax, m, nl, 1.0)
            best = max(best, -sn) if np.isfinite(sn) else best
        np.testing.assert_allclose(out["tlssn"][k, 0], best, rtol=2e-4)


def test_region_a_index_quirk_is_reproduced():
    """Region A sums the flux at scoop index i while testing the time at i+nSx
    (gtls_simple.h:214-221): a flux change before the window start moves S/N."""
    t, x = _series(seed=1, dip=(100, 3, 0.05))
    lc, tu = x[:, None].copy(), t[:, None]
    kw = dict(nw=4, nl=6, nsc=31, wmax=12.0, dw=1.0, deltat=1.0)
    base = og.gtls(lc, tu, **kw)["tlssn"][92:101, 0]
    lc2 = lc.copy()
    lc2[85, 0] += 0.3  # time 85: outside [-W/2, W/2] of every epoch 92..100
    moved = og.gtls(lc2, tu, **kw)["tlssn"][92:101, 0]
    assert np.any(moved != base)


def test_masked_samples_and_fill_factor():
    t, x = _series(seed=2, dip=(60, 2, 0.05))
    tu = t.copy()
    tu[55:66] = -1000.0  # the window around epoch 60 is mostly masked
    out = og.gtls(x[:, None], tu[:, None], 1, 5, 21, 10.0, 0.0, 1.0)
    # fill factor = valid/(W/deltat) < 0.75 -> no trial, S/N 0, W = wmax, L = H = 0
    assert out["tlssn"][60, 0] == 0.0 and out["tlsl"][60, 0] == 0.0
    assert out["tlsw"][60, 0] == np.float32(10.0)
    assert out["tlst0"][60, 0] == np.float32(60.0)


def test_absolute_mode_uses_abs_h():
    t, x = _series(seed=3, dip=(80, 3, -0.08))  # negative dip: normal mode S/N < 0
    kw = dict(nw=2, nl=5, nsc=31, wmax=10.0, dw=2.0, deltat=1.0)
    norm = og.gtls(x[:, None], t[:, None], **kw)
    ab = og.gtls(x[:, None], t[:, None], mode="absolute", **kw)
    assert ab["tlssn"][80, 0] > 0 and ab["tlshmax"][80, 0] < 0
    assert ab["tlssn"][80, 0] > norm["tlssn"][80, 0]
