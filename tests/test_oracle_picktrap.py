"""Pins oracle.picktrap (gtrap/picktrap.py:27-139 restated)."""
import numpy as np
import pytest
from scipy.interpolate import interp1d

from oracle import picktrap as opk


def test_resampling_equals_scipy_interp1d():
    rng = np.random.default_rng(0)
    for nl, nx in ((150, 128), (433, 51), (1000, 2001)):
        alpha = 2.0
        tt = np.arange(nl) * 2.0 * (1.5 * alpha) / nl - 1.5 * alpha
        tx = np.arange(nx) * 2.0 * alpha / nx - alpha
        y = rng.standard_normal(nl)
        g = (rng.random(nl) < 0.1).astype(np.int32)
        want_y = interp1d(tt, y)(tx)
        want_g = interp1d(tt, g)(tx)
        got_y = np.array([opk.np_interp_value(x, tt, y) for x in tx])
        got_g = np.array([opk.searchsorted_linear(x, tt, g) for x in tx])
        np.testing.assert_array_equal(got_y, want_y)
        np.testing.assert_array_equal(got_g, want_g)


def _curve(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    tu = (np.arange(n) * 0.0204).astype(np.float32)
    lc = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    return lc, tu


def test_window_and_clean():
    lc, tu = _curve()
    tu[1000:1003] = -1000.0   # a 3-slot gap near the candidate
    lcsx, tx, igapx, prec, st = opk.cutout(lc, tu, T0=20.5, W=0.3, nx=64)
    assert st == 0 and prec
    assert np.isclose(np.mean(lcsx), 0.0, atol=0.3)
    assert igapx.max() > 0  # the gap is visible in the mask
    # the window is centred on the first valid time >= T0
    ii = int(np.nonzero(tu.astype(np.float64) >= 20.5)[0][0])
    assert abs(tu[ii] - 20.5) < 0.0204


def test_gap_cleaning_rules():
    # one interior gap between two valid slots takes their mean; a gap with a
    # single valid neighbour takes the window median
    lc = np.array([1.0, 2.0, 9.0, 4.0, 5.0, 6.0, 7.0, 8.0], np.float32)
    tu = np.array([0.5, 1.0, -1000.0, 2.0, 2.5, 3.0, 3.5, 4.0], np.float32)
    out = opk.cutout(lc, tu, T0=2.0, W=1.0, alpha=1, nx=4, daytopix=3)
    assert out[4] == 0
    # a window that is all gaps cannot be cleaned: flagged, not looped for ever
    tu2 = np.full(20, -1000.0, np.float32)
    tu2[10] = 5.0
    tu2[9] = 0.0  # zero time: neither a gap nor a valid neighbour
    tu2[11] = 0.0
    st = opk.cutout(np.ones(20, np.float32), tu2, T0=5.0, W=1.0, alpha=1, nx=4, daytopix=3)[4]
    assert st == 2


def test_precheck_run_length():
    lc, tu = _curve(400, seed=2)
    tu[:150] = -1000.0  # long leading gap: window starts inside it
    prec = opk.cutout(lc, tu, T0=3.5, W=0.3, contcrit=20)[3]
    assert prec is False
    prec = opk.cutout(lc, tu, T0=3.5, W=0.3, contcrit=500)[3]
    assert prec is True


def test_no_valid_time_after_t0():
    lc, tu = _curve(100)
    assert opk.cutout(lc, tu, T0=1e6, W=0.3)[4] == 1
