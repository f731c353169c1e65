"""Pins oracle.peakstats (gtrap/detect_peaks.py and the statistics block of
gtrap/gtls.py:244-275 restated)."""
import numpy as np
import pytest
from scipy.signal import find_peaks

from oracle import peakstats as op


def test_docstring_cases():
    # detect_peaks.py docstring: x = [0, 1, 0, 2, 0, 3, 0, 2, 0, 1, 0] with
    # mpd = 2 keeps the peaks at 1, 5, 9 (3 suppresses its +-2 neighbours)
    x = [0, 1, 0, 2, 0, 3, 0, 2, 0, 1, 0]
    np.testing.assert_array_equal(op.detect_peaks(x, 2), [1, 5, 9])
    np.testing.assert_array_equal(op.detect_peaks(x, 1), [1, 3, 5, 7, 9])
    # flat peak, rising edge only: [0, 1, 1, 0, 1, 1, 0] -> 1 and 4
    np.testing.assert_array_equal(op.detect_peaks([0, 1, 1, 0, 1, 1, 0]), [1, 4])


@pytest.mark.parametrize("mpd", [2, 10, 48])
def test_matches_scipy_find_peaks_without_ties(mpd):
    rng = np.random.default_rng(mpd)
    for _ in range(5):
        x = rng.standard_normal(3000).cumsum() * 0.1 + rng.standard_normal(3000)
        got = op.detect_peaks(x, mpd)
        want, _ = find_peaks(x, distance=mpd + 1)
        np.testing.assert_array_equal(got, want)


def test_suppression_fixed_point():
    """Every kept peak has no kept higher-priority peak within mpd, and every
    removed candidate has one (the greedy result's defining property)."""
    rng = np.random.default_rng(3)
    x = np.round(rng.standard_normal(2000), 1)  # many exact ties
    mpd = 7
    cand = op.candidates(x)
    keep = set(op.detect_peaks(x, mpd).tolist())
    rank = {p: r for r, p in enumerate(op.priority_order(x, cand))}
    for p in cand:
        higher_kept = [q for q in keep if q != p and abs(q - p) <= mpd and rank[q] < rank[p]]
        assert (p in keep) == (len(higher_kept) == 0)


def test_nan_handling():
    x = np.array([0, 1, 0, 2, np.nan, 3, 0, 4, 0], float)
    # peaks next to the NaN (3 and 5) are dropped
    np.testing.assert_array_equal(op.detect_peaks(x, 1), [1, 7])


def test_statistics_block():
    rng = np.random.default_rng(8)
    x = np.abs(rng.standard_normal(500)) * (rng.random(500) < 0.8)
    t0 = np.arange(500) * 0.5
    s = op.peakstats(x, t0, mpd=10)
    pos = x[x > 0]
    assert s["std"] == pytest.approx(np.std(pos), rel=1e-15)
    assert s["median"] == np.median(x) and s["maxsn"] == x.max()
    assert s["far"] == pytest.approx((x.max() - np.median(x)) / np.std(pos), rel=1e-15)
    peaks = s["top"]
    assert np.all(np.diff(x[peaks]) <= 0)
    assert s["pinterval"] == abs(t0[peaks[1]] - t0[peaks[0]])
