"""Candidate cut-outs on the GPU: gtrap/picktrap.py:27-139
(pick_Wnormalized_cleaned_lc_direct) for a batch of candidates.

``cutouts(lc, tu, cq, T0, W, alpha=2, nx=128, daytopix=48, contcrit=84)``
works on the interleaved (n, nq) device image and returns
(lcsx (ncand, nx), tx, igapx (ncand, nx), prec (ncand,) bool, status) with
status 0 for a normal cut-out (see ``gtrap_cutout`` in include/gtrap_hip.h
for the others; those rows are all ones, the reference's fallback).
``pick_Wnormalized_cleaned_lc_direct(lc, tu, T0, W, ...)`` is the one-curve
form with the reference's signature and return value (lcsx, tx, igapx, prec).
"""
from __future__ import annotations

import numpy as np

from . import _native


def cutouts(lc, tu, cq, T0, W, alpha=2, nx=128, daytopix=48, contcrit=84, stream=None):
    import torch
    _native.require_gpu()
    n, nq = int(lc.shape[0]), int(lc.shape[1])
    cq = np.atleast_1d(np.asarray(cq, np.int32))
    ncand = len(cq)
    if np.any((cq < 0) | (cq >= nq)):
        raise ValueError("cutouts: curve index out of range")
    dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(np.broadcast_to(
        np.asarray(a, dt), (ncand,)))).cuda()
    lcsx = torch.empty((ncand, int(nx)), dtype=torch.float64, device="cuda")
    igapx = torch.empty((ncand, int(nx)), dtype=torch.float64, device="cuda")
    prec = torch.empty(ncand, dtype=torch.int32, device="cuda")
    status = torch.empty(ncand, dtype=torch.int32, device="cuda")
    # the candidate arrays stay referenced until the launch is queued: a
    # temporary freed right after taking its address could be handed to the
    # next allocation and overwritten before the kernel reads it
    dq, dt0, dw = dev(cq, np.int32), dev(T0, np.float64), dev(W, np.float64)
    p = _native.ptr
    _native.call("gtrap_cutout", p(lcsx), p(igapx), p(prec), p(status), p(lc), p(tu), n, nq,
                 p(dq), p(dt0), p(dw), ncand, float(alpha), int(nx), float(daytopix),
                 int(contcrit), _native.stream_handle(stream))
    tx = np.arange(int(nx)) * 2.0 * alpha / int(nx) - alpha
    return (lcsx.cpu().numpy(), tx, igapx.cpu().numpy(), prec.cpu().numpy().astype(bool),
            status.cpu().numpy())


def pick_Wnormalized_cleaned_lc_direct(lc, tu, T0, W, alpha=2, nx=128, daytopix=48,
                                       contcrit=84):
    import torch
    lc_d = torch.from_numpy(np.ascontiguousarray(np.asarray(lc, np.float32))[:, None]).cuda()
    tu_d = torch.from_numpy(np.ascontiguousarray(np.asarray(tu, np.float32))[:, None]).cuda()
    lcsx, tx, igapx, prec, status = cutouts(lc_d, tu_d, [0], [T0], [W], alpha, nx, daytopix,
                                            contcrit)
    if status[0] == 1:
        raise IndexError("no valid time at or after T0")
    return lcsx[0], tx, igapx[0], bool(prec[0])
