"""Image assembly oracle (test infrastructure only).

Restates the array part of gtrap/read_keplerlc.py:23-57 (load_keplc) and
:123-146 (throw_kepintarray) for curves already read into memory:

* every curve gets n slots filled with ``fillv`` (flux) and ``inval`` (time);
* sample i goes to slot j = cno[i] - cno[0] when its flux is not NaN:
  flux[j] = flux_i, time[j] = t_i - t0 with t0 = t[0] (``offt="t[0]"``) or 0;
* gaps are the slots with time < 0; ntrue = n minus the gaps;
* with H = max and L = min of the curve's n flux slots (fills included),
  gap slots at even j get L and at odd j get H (zig-zag fill, :39-46);
* the image is the interleaved (n, nq) float32 array, element (j, q) at
  j*nq + q, for flux and time (:50-51).

Slots j outside [0, n) are an error here (the reference exits for j >= n and
wraps negative j silently).  For duplicate cadence numbers the later sample
wins, as in the reference's loop.

Pinned by tests/test_oracle_image.py: hand-built cases (gaps, NaN samples, a
time offset, zig-zag parity, fills counted in H/L, ntrue).
"""
from __future__ import annotations

import numpy as np


def build(curves, n, fillv=1.002, inval=-1000.0, offt="t[0]"):
    """curves: list of (t float64, flux float64, cno int64).  Returns (lc, tu)
    float32 (n, nq), ntrue uint32[nq], t0 float64[nq]."""
    nq = len(curves)
    lc = np.empty((n, nq), np.float32)
    tu = np.empty((n, nq), np.float32)
    ntrue = np.empty(nq, np.uint32)
    t0s = np.empty(nq, np.float64)
    for q, (t, f, c) in enumerate(curves):
        t = np.asarray(t, np.float64)
        f = np.asarray(f, np.float64)
        c = np.asarray(c, np.int64)
        lcn = np.full(n, fillv, np.float64)
        tun = np.full(n, inval, np.float64)
        t0 = float(t[0]) if (offt == "t[0]" and len(t)) else 0.0
        for i in range(len(c)):
            j = int(c[i] - c[0])
            if not 0 <= j < n:
                raise ValueError(f"curve {q}: cadence slot {j} outside [0, {n})")
            if f[i] == f[i]:
                lcn[j] = f[i]
                tun[j] = t[i] - t0
        gap = tun < 0
        ntrue[q] = int((~gap).sum())
        hi, lo = lcn.max(), lcn.min()
        idx = np.arange(n)
        lcn[gap & (idx % 2 == 0)] = lo
        lcn[gap & (idx % 2 == 1)] = hi
        lc[:, q] = lcn.astype(np.float32)
        tu[:, q] = tun.astype(np.float32)
        t0s[q] = t0
    return lc, tu, ntrue, t0s
