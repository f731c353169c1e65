"""Periodic trapezoid least squares oracle (fold extension of gtrap/tls.py).

Definition (shared with gtrap_amd/tlsfold.py and the HIP kernel):
for each trial period P, every usable sample (t >= 0, finite flux) is folded to
u = fmod(t, P) (float64, exact); epoch e < ceil(P/dtep) sits at t0 = e*dtep;
the sample's offset is d = u - t0, replaced by (u - P) - t0 when d >= P/2 and
by (u + P) - t0 when d < -P/2.  For each width W and L = m*(0.45*W/nl)
(m < nl; the reference's normal-mode grid, include/gtls_simple.h:69,167) the
statistic of gtrap/tls.py:44-79 (get_chi2Hmax with divx=True, sigma=1) is
evaluated on d, and S/N = -H/sqrt(res/(n-3)) (tls.py:100); a trial with
n <= 3 samples in [-W/2, W/2) has no degrees of freedom and is NaN, and so is
a trial with W >= P (its window would hold two images of a sample).  The best trial of
a period is the first maximum in (W, L, epoch) order, NaN skipped
(tls.py:82-110, nanargmax).

* :func:`tlsfold` -- C, direct per-trial sums over the window samples.
* :func:`trial` -- C, one trial from the raw arrays (no candidate selection).
* :func:`trial_res` -- as :func:`trial`, also the residual and window mean.
* :func:`trial_numpy` -- the tls.py masks evaluated with numpy on folded d.
* :func:`tlsstats` -- periodogram statistics of one curve (best period, SDE).

``stat="dchi2"`` ranks and reports the detection statistic instead of S/N
(no counterpart in the reference): the least-squares fit of the trapezoid
shape g of tls.py:2-15 (unit depth) to y = x - mean over the window,
H* = sum(y g)/sum(g^2), ranked by the signed chi^2 improvement
sign(-H*) sum(y g)^2/sum(g^2) and reported as D = sign(-H*) sqrt(improvement)
/ sigma with sigma the standard deviation of the curve's usable fluxes (a
matched-filter signal-to-noise ratio); the reported H is H*.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._clib import fptr, lib


STATS = {"snr": 0, "dchi2": 1}


def tlsfold(t, x, periods, widths, nl, dtep, stat="snr"):
    t = np.ascontiguousarray(t, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    periods = np.ascontiguousarray(periods, np.float64)
    widths = np.ascontiguousarray(widths, np.float64)
    np_ = len(periods)
    sn = np.empty(np_, np.float64)
    bw, bl, be = (np.empty(np_, np.int32) for _ in range(3))
    h = np.empty(np_, np.float64)
    f = lib().orc_tlsfold
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                  C.c_int, C.c_double, C.c_int] + [C.c_void_p] * 5
    f(fptr(t), fptr(x), len(t), fptr(periods), np_, fptr(widths), len(widths), int(nl),
      float(dtep), STATS[stat], fptr(sn), fptr(bw), fptr(bl), fptr(be), fptr(h))
    return dict(sn=sn, w=bw, l=bl, e=be, h=h)


def _trial(t, x, P, e, W, m, nl, dtep, divx, stat="snr"):
    t = np.ascontiguousarray(t, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    f = lib().orc_tlsfold_trial
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                  C.c_int, C.c_double, C.c_int, C.c_int] + [C.c_void_p] * 4
    H = np.zeros(1)
    cnt = np.zeros(1, np.int32)
    res = np.zeros(1)
    mean = np.zeros(1)
    sn = f(fptr(t), fptr(x), len(t), float(P), int(e), float(W), int(m), int(nl), float(dtep),
           1 if divx else 0, STATS[stat], fptr(H), fptr(cnt), fptr(res), fptr(mean))
    return sn, H[0], int(cnt[0]), res[0], mean[0]


def trial(t, x, P, e, W, m, nl, dtep, divx=True, stat="snr"):
    """One trial: (value of the statistic, H, n)."""
    return _trial(t, x, P, e, W, m, nl, dtep, divx, stat)[:3]


def trial_res(t, x, P, e, W, m, nl, dtep, divx=True, stat="snr"):
    """One trial: (S/N, H, n, res, mean) with res the residual sum of squares
    of tls.py:74-77 and mean the window mean of tls.py:57-59 (0 without divx)."""
    return _trial(t, x, P, e, W, m, nl, dtep, divx, stat)


def fold_offsets(t, P, e, dtep, x=None):
    """Offsets d of the usable samples (t >= 0, and finite flux when x is
    given) from epoch e of period P; returns (d, mask of the usable samples)."""
    t = np.asarray(t)
    ok = t >= 0
    if x is not None:
        ok &= np.isfinite(np.asarray(x))
    u = np.fmod(t[ok].astype(np.float64), P)
    t0 = float(e) * dtep
    d = u - t0
    d = np.where(d >= 0.5 * P, (u - P) - t0, np.where(d < -0.5 * P, (u + P) - t0, d))
    return d, ok


def chi2hmax_numpy(tn, xx, W, L, divx=True, sigma=1.0):
    """get_chi2Hmax of tls.py:44-79 on offsets tn (float64 numpy, nan-aware).
    Returns (Hmax, res, n, xsum)."""
    tn = np.asarray(tn, np.float64)
    xx = np.asarray(xx, np.float64)
    maskA = (tn >= -W / 2) * (tn < -W / 4 - L / 2)
    maskB = (tn >= -W / 4 - L / 2) * (tn < -W / 4 + L / 2)
    maskC = (tn < W / 4 - L / 2) * (tn >= -W / 4 + L / 2)
    maskD = (tn < W / 4 + L / 2) * (tn >= W / 4 - L / 2)
    maskE = (tn < W / 2) * (tn >= W / 4 + L / 2)
    nA, nB, nC, nD, nE = (int(k.sum()) for k in (maskA, maskB, maskC, maskD, maskE))
    n = nA + nB + nC + nD + nE
    with np.errstate(all="ignore"):
        xsum = np.nanmean(xx[(tn >= -W / 2) * (tn <= W / 2)]) if divx else 0.0
        x_ = xx - xsum
        xAEC = np.nansum(x_[maskA]) + np.nansum(x_[maskE]) + np.nansum(x_[maskC])
        xBD = np.nansum(x_[maskB]) + np.nansum(x_[maskD])
        xtB = np.nansum(x_[maskB] * tn[maskB])
        xtD = np.nansum(x_[maskD] * tn[maskD])
        tB = np.nansum(tn[maskB])
        tD = np.nansum(tn[maskD])
        t2BD = np.nansum(tn[maskB] * tn[maskB]) + np.nansum(tn[maskD] * tn[maskD])
        denA = 4 * L * L * (nA - nC + nE) + W * W * (nB + nD) + 8 * W * (tB - tD) + 16 * t2BD
        numB = 8 * L * L * xAEC - 4 * W * L * xBD + 16 * L * (xtD - xtB)
        H = -numB / denA
        Hh = H / 2.0
        res = (np.nansum((x_[maskA] + Hh) ** 2) + np.nansum((x_[maskC] - Hh) ** 2)
               + np.nansum((x_[maskE] + Hh) ** 2))
        res = res + np.nansum((x_[maskB] - H / L * (tn[maskB] + W / 4.0)) ** 2)
        res = res + np.nansum((x_[maskD] + H / L * (tn[maskD] - W / 4.0)) ** 2)
        res = res / (sigma * sigma)
    return H, res, n, xsum


def trial_numpy(t, x, P, e, W, m, nl, dtep, divx=True):
    """tls.py's get_chi2Hmax masks and formulas (tls.py:44-79) on folded d;
    returns (S/N, H, n) with S/N = -H/sqrt(res/(n-3)) (tls.py:100)."""
    d, ok = fold_offsets(t, P, e, dtep, x)
    xx = np.asarray(x)[ok].astype(np.float64)
    L = m * (0.45 * W / nl)
    H, res, n, _ = chi2hmax_numpy(d, xx, W, L, divx=divx)
    with np.errstate(all="ignore"):
        sn = -H / np.sqrt(res / (n - 3)) if n > 3 else np.nan
    return sn, H, n


def tlsstats(sn):
    """Periodogram statistics of one curve's per-period S/N (float64).

    best: the first maximum over the values that are neither NaN nor -inf
    (-inf marks a period without any trial; +inf, a perfect fit, can win),
    -1 when there is none; mean and variance over the finite values;
    SDE = (best value - mean) / sqrt(var), the statistic of
    gtrap/getstat.py:114-118 (ave = sum/n, var = sumsq/n - ave^2) applied to
    the trapezoid periodogram.  Returns (best index, best value, SDE)."""
    sn = np.asarray(sn, np.float64)
    cand = ~np.isnan(sn) & ~np.isneginf(sn)
    if not cand.any():
        best, bval = -1, -np.inf
    else:
        v = np.where(cand, sn, -np.inf)
        best = int(np.argmax(v))
        bval = float(v[best])
    fin = sn[np.isfinite(sn)]
    with np.errstate(all="ignore"):
        mean = fin.sum() / len(fin) if len(fin) else np.nan
        var = (fin * fin).sum() / len(fin) - mean * mean if len(fin) else np.nan
        sde = (bval - mean) / np.sqrt(var)
    return best, bval, float(sde)


def period_numpy(t, x, P, widths, nl, dtep):
    """Best trial of one period with numpy, tls.py's statistic vectorised over
    the epochs (the CPU path the benchmark times beside the GPU): every
    usable sample is folded (:func:`fold_offsets`' rule), each epoch gathers
    the samples within the widest half window (exact offsets decide every
    region), and get_chi2Hmax (tls.py:44-79) runs on (epoch, sample) arrays
    for each (W, L).  Returns (S/N, w, m, e, H) of the first maximum in
    (W, L, epoch) order, NaN skipped (tls.py:100-102); (-inf, -1, -1, -1, NaN)
    when no trial exists."""
    t = np.asarray(t)
    x = np.asarray(x)
    ok = (t >= 0) & np.isfinite(x)
    u = np.fmod(t[ok].astype(np.float64), P)
    xv = x[ok].astype(np.float64)
    o = np.argsort(u, kind="stable")
    u, xv = u[o], xv[o]
    k = len(u)
    nE = int(np.ceil(P / dtep))
    t0 = np.arange(nE) * dtep
    hw = 0.5 * float(np.max(widths)) + 1e-9
    if k == 0:
        return -np.inf, -1, -1, -1, np.nan
    if 2.0 * hw < P:
        ue = np.concatenate([u - P, u, u + P])
        src = np.tile(np.arange(k), 3)
        lo = np.searchsorted(ue, t0 - hw, "left")
        hi = np.searchsorted(ue, t0 + hw, "right")
        K = max(1, int((hi - lo).max()))
        idx = lo[:, None] + np.arange(K)[None, :]
        valid = idx < hi[:, None]
        idx = src[np.minimum(idx, 3 * k - 1)]
    else:
        idx = np.broadcast_to(np.arange(k), (nE, k))
        valid = np.ones((nE, k), bool)
    uu = u[idx]
    c = t0[:, None]
    d = uu - c
    d = np.where(d >= 0.5 * P, (uu - P) - c, np.where(d < -0.5 * P, (uu + P) - c, d))
    xx = np.where(valid, xv[idx], 0.0)
    best = (-np.inf, -1, -1, -1, np.nan)
    with np.errstate(all="ignore"):
        for w, W in enumerate(widths):
            if not W < P:
                continue
            close = valid & (d >= -W / 2) & (d <= W / 2)
            mean = (xx * close).sum(1) / close.sum(1)
            y = xx - mean[:, None]
            for m in range(1, nl):
                L = m * (0.45 * W / nl)
                mA = valid & (d >= -W / 2) & (d < -W / 4 - L / 2)
                mB = valid & (d >= -W / 4 - L / 2) & (d < -W / 4 + L / 2)
                mC = valid & (d < W / 4 - L / 2) & (d >= -W / 4 + L / 2)
                mD = valid & (d < W / 4 + L / 2) & (d >= W / 4 - L / 2)
                mE = valid & (d < W / 2) & (d >= W / 4 + L / 2)
                nA, nB, nC, nD, nE_ = (mk.sum(1) for mk in (mA, mB, mC, mD, mE))
                n = nA + nB + nC + nD + nE_
                xAEC = (y * mA).sum(1) + (y * mE).sum(1) + (y * mC).sum(1)
                xBD = (y * mB).sum(1) + (y * mD).sum(1)
                xtB = (y * d * mB).sum(1)
                xtD = (y * d * mD).sum(1)
                tB = (d * mB).sum(1)
                tD = (d * mD).sum(1)
                t2BD = (d * d * mB).sum(1) + (d * d * mD).sum(1)
                denA = 4 * L * L * (nA - nC + nE_) + W * W * (nB + nD) + 8 * W * (tB - tD) + \
                    16 * t2BD
                numB = 8 * L * L * xAEC - 4 * W * L * xBD + 16 * L * (xtD - xtB)
                H = -numB / denA
                Hh = H / 2.0
                res = ((((y + Hh[:, None]) ** 2) * (mA | mE)).sum(1)
                       + (((y - Hh[:, None]) ** 2) * mC).sum(1)
                       + (((y - (H / L)[:, None] * (d + W / 4.0)) ** 2) * mB).sum(1)
                       + (((y + (H / L)[:, None] * (d - W / 4.0)) ** 2) * mD).sum(1))
                sn = np.where(n > 3, -H / np.sqrt(res / (n - 3)), np.nan)
                if np.all(np.isnan(sn)):
                    continue
                e = int(np.nanargmax(sn))
                if sn[e] > best[0]:
                    best = (float(sn[e]), w, m, e, float(H[e]))
    return best
