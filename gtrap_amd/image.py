"""Device image assembly: ragged light curves -> the interleaved (n, nq) image.

Drop-in for the array part of gtrap/read_keplerlc.py:load_keplc
(:23-57, with throw_kepintarray :123-146): the caller has already read the
curves (time, normalised flux, cadence numbers; FITS parsing is outside this
package).  Each curve is placed by cadence number into n slots, missing or
NaN samples become gaps (time ``inval``), gaps are zig-zag filled with the
curve's minimum (even slots) and maximum (odd slots), and ``ntrue`` counts
the valid slots.  The work runs on the GPU (``gtrap_image_build``).
"""
from __future__ import annotations

import numpy as np

from . import _native


def build_device(t, flux, cno, off, length, nq, n, maxlen, fillv=1.002, inval=-1000.0,
                 offt=True, stream=None):
    """Device inputs (flat float64 t and flux, int64 cno, int64 off[nq],
    int32 len[nq]) -> dict of device tensors lc, tu (n, nq) float32, ntrue
    uint32 (as int32), hfill, lfill float32, t0 float64, nbad int32."""
    import torch
    _native.require_gpu()
    dev = "cuda"
    out = dict(lc=torch.empty((n, nq), dtype=torch.float32, device=dev),
               tu=torch.empty((n, nq), dtype=torch.float32, device=dev),
               ntrue=torch.empty(nq, dtype=torch.int32, device=dev),
               hfill=torch.empty(nq, dtype=torch.float32, device=dev),
               lfill=torch.empty(nq, dtype=torch.float32, device=dev),
               t0=torch.empty(nq, dtype=torch.float64, device=dev),
               nbad=torch.empty(1, dtype=torch.int32, device=dev))
    p = _native.ptr
    _native.call("gtrap_image_build", p(out["lc"]), p(out["tu"]), p(out["ntrue"]),
                 p(out["hfill"]), p(out["lfill"]), p(out["t0"]), p(t), p(flux), p(cno), p(off),
                 p(length), int(nq), int(n), int(maxlen), float(fillv), float(inval),
                 1 if offt else 0, p(out["nbad"]), _native.stream_handle(stream))
    return out


def load_curves(curves, n=72000, fillv=1.002, inval=-1000.0, offt="t[0]"):
    """``curves``: list of (t, flux, cadence number) host arrays, one per curve.
    Returns (lc, tu, n, ntrue, nq, inval, t0arr) as load_keplc does (device
    tensors for lc, tu and ntrue, numpy for t0arr).  A cadence slot outside
    [0, n) raises ValueError, as the reference stops there."""
    import torch
    nq = len(curves)
    if nq == 0:
        raise ValueError("load_curves: no curves")
    lens = np.array([len(c[0]) for c in curves], np.int64)
    if np.any(lens == 0):
        raise ValueError("load_curves: empty curve")
    for k, (t, f, c) in enumerate(curves):
        if not (len(t) == len(f) == len(c)):
            raise ValueError(f"load_curves: curve {k}: t, flux, cno lengths differ")
    off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    cat = lambda i, dt: torch.from_numpy(
        np.ascontiguousarray(np.concatenate([np.asarray(c[i], dt) for c in curves]))).cuda()
    out = build_device(cat(0, np.float64), cat(1, np.float64), cat(2, np.int64),
                       torch.from_numpy(off).cuda(),
                       torch.from_numpy(lens.astype(np.int32)).cuda(), nq, n, int(lens.max()),
                       fillv, inval, offt == "t[0]")
    nbad = int(out["nbad"].item())
    if nbad:
        raise ValueError(f"load_curves: {nbad} samples fall outside the {n} cadence slots; "
                         "set a larger n")
    return (out["lc"], out["tu"], n, out["ntrue"], nq, inval, out["t0"].cpu().numpy())
