"""Pins oracle.image.build (gtrap/read_keplerlc.py:23-57,123-146 restated)
with hand-built cases."""
import numpy as np
import pytest

from oracle import image as oi


def test_hand_case():
    # cadences 100..106 with 102 missing, a NaN flux at 104, t0 = 10.0
    t = np.array([10.0, 10.5, 11.5, 12.0, 12.5, 13.0])
    f = np.array([1.0, 0.99, 1.01, np.nan, 0.98, 1.03])
    c = np.array([100, 101, 103, 104, 105, 106])
    lc, tu, ntrue, t0 = oi.build([(t, f, c)], n=10)
    assert t0[0] == 10.0
    np.testing.assert_array_equal(tu[:, 0], np.float32([0, .5, -1000, 1.5, -1000, 2.5, 3.0,
                                                        -1000, -1000, -1000]))
    assert ntrue[0] == 5
    # H/L over the 10 slots include the 1.002 fills: H = 1.03, L = 0.98
    hi, lo = np.float32(1.03), np.float32(0.98)
    want = np.float32([1.0, 0.99, lo, 1.01, lo, 0.98, 1.03, hi, lo, hi])
    np.testing.assert_array_equal(lc[:, 0], want)


def test_fill_value_enters_extrema_and_no_offset():
    t = np.array([5.0, 6.0])
    f = np.array([0.5, 0.6])
    c = np.array([0, 3])
    lc, tu, ntrue, _ = oi.build([(t, f, c)], n=5, offt="0")
    np.testing.assert_array_equal(tu[:, 0], np.float32([5.0, -1000, -1000, 6.0, -1000]))
    # max over the slots is the 1.002 fill itself, min the 0.5 sample
    np.testing.assert_array_equal(lc[:, 0], np.float32([0.5, 1.002, 0.5, 0.6, 0.5]))
    assert ntrue[0] == 2


def test_interleaved_layout_and_errors():
    a = (np.array([0.0, 1.0]), np.array([1.0, 2.0]), np.array([7, 8]))
    b = (np.array([3.0]), np.array([4.0]), np.array([1]))
    lc, tu, ntrue, _ = oi.build([a, b], n=3)
    assert lc.shape == (3, 2) and lc[1, 0] == 2.0 and lc[0, 1] == 4.0
    np.testing.assert_array_equal(ntrue, [2, 1])
    with pytest.raises(ValueError):
        oi.build([(np.array([0.0, 1.0]), np.array([1.0, 1.0]), np.array([5, 9]))], n=3)
