"""Peak statistics oracle (test infrastructure only).

Restates the post-search statistics of the reference's single-transit
drivers for one S/N series x (gtrap/gtls.py:244-275,
kepler/gtls_mockkepler.py:336-376) and the peak finder they call
(gtrap/detect_peaks.py:10-138, edge='rising', no height or threshold
filter, kpsh False):

* candidates: 1 <= i <= n-2 with x[i] > x[i-1] and x[i+1] <= x[i], none of
  x[i-1], x[i], x[i+1] NaN (detect_peaks.py:96-123);
* with mpd > 1, candidates in descending height (ties: the larger index
  first, the order a stable sort gives before the reversal) are taken
  greedily; each kept peak removes every candidate within mpd samples
  (detect_peaks.py:124-131);
* peaks sorted by descending height; ntop = 4; maxsn = max(x);
  median = median(x); std = population std of x[x > 0]; far = (maxsn -
  median) / std; with more than ntop peaks diff = (x[peak[ntop-1]] -
  median) / std and pssn = (x[peak[0]] - median of x[peak[ntop:]]) / std of
  x[peak[ntop:]], else diff = 0 and pssn = inf; pinterval = |t0[peak[1]] -
  t0[peak[0]]| (NaN with fewer than two peaks); lent = count(x > 0).

Everything in float64 (the reference's numpy runs in float32, so its
statistics agree with these to float32 rounding).  Ties in height beyond 16
candidates are ordered by numpy's unstable sort in the reference: parity
unpinned there.

Pinned by tests/test_oracle_peakstats.py: the two hand cases of
detect_peaks.py's docstring, scipy.signal.find_peaks(distance=mpd+1) on
tie-free series, a brute-force fixed-point check of the suppression, NaN
handling, and numpy's median/std for the statistics.
"""
from __future__ import annotations

import numpy as np


def candidates(x):
    x = np.asarray(x, np.float64)
    n = len(x)
    if n < 3:
        return np.array([], np.int64)
    i = np.arange(1, n - 1)
    a, b, c = x[i - 1], x[i], x[i + 1]
    ok = ~(np.isnan(a) | np.isnan(b) | np.isnan(c)) & (b > a) & (c <= b)
    return i[ok]


def priority_order(x, ind):
    """Candidates by descending height, ties: larger index first."""
    x = np.asarray(x, np.float64)
    return ind[np.lexsort((-ind, -x[ind]))]


def detect_peaks(x, mpd=1):
    ind = candidates(x)
    if ind.size and mpd > 1:
        order = priority_order(x, ind)
        covered = np.zeros(len(x) + 2 * mpd + 1, bool)
        keep = []
        for p in order:
            if not covered[p + mpd]:
                keep.append(p)
                covered[p:p + 2 * mpd + 1] = True
        ind = np.sort(np.array(keep, np.int64))
    return ind


def peakstats(x, t0=None, mpd=10, ntop=4):
    x = np.asarray(x, np.float64)
    pos = x[x > 0]
    std = pos.std() if pos.size else np.nan
    median = float(np.median(x))
    maxsn = float(np.max(x))
    peaks = detect_peaks(x, mpd)
    peaks = priority_order(x, peaks)
    out = dict(maxsn=maxsn, median=median, std=float(std), lent=int((x > 0).sum()),
               npeak=int(len(peaks)), far=(maxsn - median) / std if std else np.nan,
               top=peaks[:ntop].copy())
    if len(peaks) > ntop:
        tail = x[peaks[ntop:]]
        out["diff"] = (x[peaks[ntop - 1]] - median) / std
        out["pssn"] = (x[peaks[0]] - np.nanmedian(tail)) / np.nanstd(tail)
    else:
        out["diff"] = 0.0
        out["pssn"] = np.inf
    out["pinterval"] = (abs(float(t0[peaks[1]]) - float(t0[peaks[0]]))
                        if (t0 is not None and len(peaks) >= 2) else np.nan)
    return out
