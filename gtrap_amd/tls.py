"""Trapezoid least squares: drop-in for gtrap/tls.py with the grid on the GPU.

* ``gen_trapzoid(t, H, W, L, t0)`` -- the trapezoid model (tls.py:2-15):
  -H/2 outside, H/2 on the flat bottom, linear ingress/egress of width L
  centred at -W/4 and W/4.
* ``get_chi2Hmax(t, xx, W, L, t0, sigma=1.0, divx=True)`` -> (Hmax, res, n, xsum)
  (tls.py:44-79) and ``get_Hmax(t, x, W, L, t0)`` (tls.py:17-42, no mean
  removal), evaluated by the ``tlsgrid_kernel``.
* ``tlsfit(t, x, Wl, Ll, tl)`` -> (xpre, H, W, L, T0, offset) (tls.py:82-110):
  every (W, L, t0) trial on the GPU, S/N = -H/sqrt(res/(n-3)), first maximum
  in (W, L, t0) order with NaN skipped (numpy's nanargmax).
"""
from __future__ import annotations

import numpy as np

from . import _native


def gen_trapzoid(t, H, W, L, t0):
    tn = np.asarray(t, dtype=np.float64) - t0
    x = np.zeros(len(tn))
    x[tn <= -W / 4 - L / 2] = -H / 2
    x[tn >= W / 4 + L / 2] = -H / 2
    x[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = H / 2
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    x[mB] = H / L * (tn[mB] + W / 4)
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    x[mD] = -H / L * (tn[mD] - W / 4)
    return x


def grid_device(t, x, Wl, Ll, tl, sigma=1.0, divx=True, stream=None):
    """All (W, L, t0) trials: returns device tensors H, res, n, xsum of shape
    (len(Wl), len(Ll), len(tl))."""
    import torch
    _native.require_gpu()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64).reshape(-1)).cuda()
    t_, x_, W_, L_, T_ = dev(t), dev(x), dev(Wl), dev(Ll), dev(tl)
    shape = (W_.numel(), L_.numel(), T_.numel())
    ntr = int(np.prod(shape))
    H = torch.empty(ntr, dtype=torch.float64, device="cuda")
    res = torch.empty(ntr, dtype=torch.float64, device="cuda")
    nn = torch.empty(ntr, dtype=torch.int32, device="cuda")
    xs = torch.empty(ntr, dtype=torch.float64, device="cuda")
    p = _native.ptr
    tt = np.asarray(t, np.float64).reshape(-1)
    srt = bool(tt.size < 2 or np.all(tt[1:] >= tt[:-1]))  # NaN compares false: unsorted
    _native.call("gtrap_tlsgrid", p(t_), p(x_), int(t_.numel()), p(W_), shape[0], p(L_),
                 shape[1], p(T_), shape[2], 1 if divx else 0, float(sigma), int(srt), p(H),
                 p(res), p(nn), p(xs), _native.stream_handle(stream))
    return (H.view(shape), res.view(shape), nn.view(shape), xs.view(shape))


def get_chi2Hmax(t, xx, W, L, t0, sigma=1.0, divx=True):
    H, res, n, xs = grid_device(t, xx, [W], [L], [t0], sigma=sigma, divx=divx)
    return float(H.item()), float(res.item()), int(n.item()), float(xs.item())


def get_Hmax(t, x, W, L, t0):
    return get_chi2Hmax(t, x, W, L, t0, divx=False)[0]


def tlsfit(t, x, Wl, Ll, tl):
    H, res, n, xs = grid_device(t, x, Wl, Ll, tl)
    Hh, rh, nh, xh = (a.cpu().numpy().reshape(-1) for a in (H, res, n, xs))
    with np.errstate(all="ignore"):
        sn = -Hh / np.sqrt(rh / (nh - 3))
    imax = int(np.nanargmax(sn))
    nL, nT = len(Ll), len(tl)
    w, l, k = imax // (nL * nT), (imax // nT) % nL, imax % nT
    Hb, Wb, Lb, T0, off = Hh[imax], Wl[w], Ll[l], tl[k], xh[imax]
    xpre = gen_trapzoid(t, Hb, Wb, Lb, T0) + off
    return xpre, Hb, Wb, Lb, T0, off
