"""Pins the median-detrend oracle (gtrap/gfilter.py:13-85).

The reference's bisection selector (restated in C) must equal scipy's
zero-padded medfilt and a brute-force np.sort median, bit for bit.
"""
import numpy as np
import pytest

from oracle import medfilt as om


def _cases():
    rng = np.random.default_rng(7)
    yield "gauss", (1.0 + 1e-3 * rng.standard_normal((257, 5))).astype(np.float32)
    yield "ties", rng.integers(-3, 4, size=(200, 4)).astype(np.float32)
    yield "const", np.full((64, 3), 1.25, np.float32)
    yield "tiny", (1.0 + rng.standard_normal((7, 2))).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal((300, 3))).astype(np.float32)
    x[rng.random(x.shape) < 0.02] += 1e-2
    yield "outliers", x
    # adjacent floats: the bisection midpoint rounds onto an endpoint
    base = np.float32(1.0)
    adj = np.array([base, np.nextafter(base, np.float32(2)), np.nextafter(base, np.float32(0))],
                   np.float32)
    yield "adjacent", adj[rng.integers(0, 3, size=(150, 2))]
    yield "negative", (-1.0 + 1e-2 * rng.standard_normal((128, 2))).astype(np.float32)


@pytest.mark.parametrize("r", [0, 1, 3, 7, 16, 24])
@pytest.mark.parametrize("mode", ["sub", "div"])
def test_bisection_equals_scipy_and_sort(r, mode):
    for name, img in _cases():
        a = om.gfilter_bisect(img, r, mode)
        b = om.detrend_scipy(img, r, mode)
        c = om.detrend_sort(img, r, mode)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=name)
        np.testing.assert_array_equal(b.view(np.uint32), c.view(np.uint32), err_msg=name)


def test_invariants():
    # constant curve: interior median equals the value -> sub gives 0, div gives 1
    img = np.full((100, 2), 3.0, np.float32)
    s = om.detrend_scipy(img, 5, "sub")
    d = om.detrend_scipy(img, 5, "div")
    assert np.all(s[5:-5] == 0.0) and np.all(d[5:-5] == 1.0)
    # edges see zero padding: with r >= n/2 on a positive curve the first sample's
    # window holds r zeros and r+1 values -> the median is the smallest value
    img = np.arange(1, 6, dtype=np.float32)[:, None]
    med = om.median_scipy(img, 2)
    assert med[0, 0] == 1.0 and med[4, 0] == 3.0
