"""Candidate cut-out oracle (test infrastructure only).

Restates gtrap/picktrap.py:27-139 (pick_Wnormalized_cleaned_lc_direct) for
one curve (lc, tu) and a candidate (T0, W):

* ii = the first slot whose time is the first valid (tu > 0) time >= T0,
  compared in float64 as numpy's searchsorted does (:31-35);
  wid = int(alpha*W*daytopix), h = int(1.11*alpha*wid), the window is
  slots [ii-h, ii+h) (:37-50);
* slots outside the curve get time 1000 and flux = nanmedian of the curve's
  slots inside the window (float32 arithmetic, as numpy on float32) (:53-58);
* prec (:66-77): False if the window is empty or, when its first time is < 0
  or its last < -1, some run of negative times is longer than contcrit;
* gap cleaning (:82-112), a Jacobi sweep repeated until no time is < 0: a
  gap slot with valid (> 0) neighbours on both sides takes their mean, with
  one valid neighbour the median of the window's current fluxes; filled
  slots get time 1000 and igap 1.  A sweep that fills nothing stops the
  cleaning (the reference would loop for ever) and the result is flagged;
* flux normalised by its mean and population std (:117);
* resampled from tt = k*2*alphad/nlcs - alphad (alphad = 1.5 alpha) onto
  tx = i*2*alpha/nx - alpha: flux as scipy's interp1d does for float64 data
  (np.interp), igap with interp1d's searchsorted formula for integer data
  (slope from the bracketing pair found by searchsorted, clipped to
  [1, nlcs-1]); a tx outside [tt[0], tt[-1]] gives all ones (:119-134).

Pinned by tests/test_oracle_picktrap.py: hand cases of each step, and the
resampling against scipy.interpolate.interp1d itself.
"""
from __future__ import annotations

import numpy as np


def np_interp_value(x, xp, fp):
    """np.interp for one x inside [xp[0], xp[-1]] (the formula numpy uses)."""
    n = len(xp)
    if x == xp[-1]:
        return float(fp[-1])
    j = int(np.searchsorted(xp, x, side="right")) - 1
    j = min(max(j, 0), n - 2)
    slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j])
    y = slope * (x - xp[j]) + fp[j]
    if np.isnan(y):
        y = slope * (x - xp[j + 1]) + fp[j + 1]
    return float(y)


def searchsorted_linear(x, xp, fp):
    """interp1d's integer-data formula for one x."""
    k = int(np.searchsorted(xp, x, side="left"))
    k = min(max(k, 1), len(xp) - 1)
    lo, hi = k - 1, k
    slope = (float(fp[hi]) - float(fp[lo])) / (xp[hi] - xp[lo])
    return float(slope * (x - xp[lo]) + float(fp[lo]))


def cutout(lc, tu, T0, W, alpha=2, nx=128, daytopix=48, contcrit=84):
    """Returns (lcsx, tx, igapx, prec, status): status 0 ok, 1 no valid time
    >= T0, 2 cleaning stuck (a gap with no valid neighbour), 3 window < 2."""
    lc = np.asarray(lc, np.float32)
    tu = np.asarray(tu, np.float32)
    n = len(tu)
    tx = np.arange(nx) * 2.0 * alpha / nx - alpha
    valid = np.nonzero(tu > 0.0)[0]
    after = valid[tu[valid].astype(np.float64) >= float(T0)] if valid.size else valid
    if after.size == 0:
        return np.ones(nx), tx, np.ones(nx), False, 1
    ii = int(np.nonzero(tu == tu[after[0]])[0][0])
    wid = int(alpha * W * daytopix)
    h = int(1.11 * alpha * wid)
    iss, iee = ii - h, ii + h
    nget = iee - iss
    if nget < 2:
        return np.ones(nx), tx, np.ones(nx), False, 3
    istart, isp = (0, -iss) if iss < 0 else (iss, 0)
    iend = min(iee, n)
    iep = iend - istart + isp
    seg = lc[istart:iend]
    seg = seg[~np.isnan(seg)]
    if seg.size == 0:
        med = np.float32(np.nan)
    else:
        s = np.sort(seg)
        m = len(s)
        med = s[m // 2] if m % 2 else np.float32((s[m // 2 - 1] + s[m // 2]) / np.float32(2))
    tus = np.full(nget, 1000.0)
    lcs = np.full(nget, float(med))
    igap = np.zeros(nget, np.int32)
    tus[isp:iep] = tu[istart:iend]
    lcs[isp:iep] = lc[istart:iend]
    prec = True
    if tus[0] < 0.0 or tus[-1] < -1:
        run = 0
        for v in tus:
            run = run + 1 if v < 0.0 else 0
            if run > contcrit:
                prec = False
    status = 0
    while np.any(tus < 0.0):
        plus_t = np.concatenate([[-2000.0], tus[:-1]])
        minus_t = np.concatenate([tus[1:], [-2000.0]])
        plus_l = np.concatenate([[0.0], lcs[:-1]])
        minus_l = np.concatenate([lcs[1:], [0.0]])
        medv = float(np.median(lcs))
        gap = tus < 0.0
        both = gap & (minus_t > 0.0) & (plus_t > 0.0)
        one = gap & ~both & ((minus_t > 0.0) | (plus_t > 0.0))
        if not (both.any() or one.any()):
            status = 2
            break
        lcs = np.where(both, (minus_l + plus_l) / 2.0, np.where(one, medv, lcs))
        igap = np.where(both | one, 1, igap).astype(np.int32)
        tus = np.where(both | one, 1000.0, tus)
    lcs = (lcs - lcs.mean()) / lcs.std()
    nl = len(lcs)
    alphad = alpha * 1.5
    tt = np.arange(nl) * 2.0 * alphad / nl - alphad
    if tx.min() < tt[0] or tx.max() > tt[-1]:
        return np.ones(nx), tx, np.ones(nx), prec, status
    lcsx = np.array([np_interp_value(x, tt, lcs) for x in tx])
    igapx = np.array([searchsorted_linear(x, tt, igap) for x in tx])
    return lcsx, tx, igapx, prec, status
