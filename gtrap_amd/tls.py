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
    return _trapzoid_masks(tn, H, W, L)


def _trapzoid_masks(tn, H, W, L):
    x = np.zeros(len(tn))
    x[tn <= -W / 4 - L / 2] = -H / 2
    x[tn >= W / 4 + L / 2] = -H / 2
    x[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = H / 2
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    x[mB] = H / L * (tn[mB] + W / 4)
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    x[mD] = -H / L * (tn[mD] - W / 4)
    return x


def _trapzoid_sorted(tn, H, W, L):
    """gen_trapzoid for non-decreasing tn (no NaN): every region of the
    reference's mask assignments is a contiguous run of the sorted times, so
    each becomes one slice (found by binary search), written in the same order
    with the same expressions; equal to the mask form bit for bit."""
    x = np.zeros(len(tn))
    ss = np.searchsorted
    a1, a2, a3, a4 = -W / 4 - L / 2, W / 4 + L / 2, W / 4 - L / 2, -W / 4 + L / 2
    x[:ss(tn, a1, "right")] = -H / 2  # tn <= -W/4 - L/2
    x[ss(tn, a2, "left"):] = -H / 2  # tn >= W/4 + L/2
    lo, hi = ss(tn, a4, "left"), ss(tn, a3, "left")  # -W/4 + L/2 <= tn < W/4 - L/2
    if hi > lo:
        x[lo:hi] = H / 2
    lo, hi = ss(tn, a1, "right"), ss(tn, a4, "right")  # -W/4 - L/2 < tn <= -W/4 + L/2
    if hi > lo:
        x[lo:hi] = H / L * (tn[lo:hi] + W / 4)
    lo, hi = ss(tn, a3, "left"), ss(tn, a2, "left")  # W/4 - L/2 <= tn < W/4 + L/2
    if hi > lo:
        x[lo:hi] = -H / L * (tn[lo:hi] - W / 4)
    return x


def _is_sorted(tt):
    return bool(tt.size < 2 or np.all(tt[1:] >= tt[:-1]))  # NaN compares false: unsorted


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
    srt = _is_sorted(np.asarray(t, np.float64).reshape(-1))
    _native.call("gtrap_tlsgrid", p(t_), p(x_), int(t_.numel()), p(W_), shape[0], p(L_),
                 shape[1], p(T_), shape[2], 1 if divx else 0, float(sigma), int(srt), p(H),
                 p(res), p(nn), p(xs), _native.stream_handle(stream))
    return (H.view(shape), res.view(shape), nn.view(shape), xs.view(shape))


def get_chi2Hmax(t, xx, W, L, t0, sigma=1.0, divx=True):
    H, res, n, xs = grid_device(t, xx, [W], [L], [t0], sigma=sigma, divx=divx)
    return float(H.item()), float(res.item()), int(n.item()), float(xs.item())


def get_Hmax(t, x, W, L, t0):
    return get_chi2Hmax(t, x, W, L, t0, divx=False)[0]


class _FitBuffers:
    """Reused buffers of tlsfit on one stream: a pinned staging area and its
    device copy for the inputs (t, x, Wl, Ll, tl), the grid's outputs, and
    the four doubles of the choice (device and pinned)."""

    def __init__(self):
        self.nin = self.ntr = -1

    def fit(self, nin, ntr):
        import torch
        if nin > self.nin:
            self.nin = max(nin, 1 << 12)
            self.host = torch.empty(self.nin, dtype=torch.float64, pin_memory=True)
            self.dev = torch.empty(self.nin, dtype=torch.float64, device="cuda")
        if ntr > self.ntr:
            self.ntr = max(ntr, 1 << 10)
            self.out = torch.empty((3, self.ntr), dtype=torch.float64, device="cuda")
            self.nn = torch.empty(self.ntr, dtype=torch.int32, device="cuda")
        if not hasattr(self, "best"):
            self.best = torch.empty(4, dtype=torch.float64, device="cuda")
            self.best_host = torch.empty(4, dtype=torch.float64, pin_memory=True)
        return self


_FIT: dict = {}


def tlsfit(t, x, Wl, Ll, tl):
    """tls.py:82-110 on the device: the grid (tlsgrid_kernel), then the S/N
    and nanargmax of tls.py:100-110 (tlsfit_best_kernel); one upload of the
    inputs and one 32-byte download of the choice per call."""
    import torch
    _native.require_gpu()
    f64 = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)
    tt, xx, W_, L_, T_ = f64(t), f64(x), f64(Wl), f64(Ll), f64(tl)
    n, nW, nL, nT = tt.size, W_.size, L_.size, T_.size
    ntr = nW * nL * nT
    if ntr == 0:
        raise ValueError("attempt to get argmax of an empty sequence")
    stream = torch.cuda.current_stream()
    key = _native.stream_handle(stream)
    buf = _FIT.setdefault(key, _FitBuffers()).fit(2 * n + nW + nL + nT, ntr)
    hv = buf.host.numpy()
    offs = np.cumsum([0, n, n, nW, nL, nT])
    for a, lo, hi in zip((tt, xx, W_, L_, T_), offs[:-1], offs[1:]):
        hv[lo:hi] = a
    buf.dev[:offs[-1]].copy_(buf.host[:offs[-1]], non_blocking=True)
    p = _native.ptr
    dp = lambda k: buf.dev.data_ptr() + 8 * int(offs[k])
    H, res, xs = (buf.out[i] for i in range(3))
    srt = _is_sorted(tt)
    _native.call("gtrap_tlsgrid", dp(0), dp(1), n, dp(2), nW, dp(3), nL, dp(4), nT, 1, 1.0,
                 int(srt), p(H), p(res), p(buf.nn), p(xs), key)
    _native.call("gtrap_tlsfit_best", p(H), p(res), p(buf.nn), p(xs), ntr, None, p(buf.best),
                 key)
    buf.best_host.copy_(buf.best, non_blocking=True)
    stream.synchronize()
    bh = buf.best_host.numpy().copy()
    if bh[0] < 0:
        raise ValueError("All-NaN slice encountered")
    imax, Hb, off = int(bh[0]), bh[2], bh[3]  # numpy float64 scalars, as the grid's
    w, l, k = imax // (nL * nT), (imax // nT) % nL, imax % nT
    Wb, Lb, T0 = Wl[w], Ll[l], tl[k]
    tn = tt - T0
    model = _trapzoid_sorted(tn, Hb, Wb, Lb) if srt and n >= 2 else _trapzoid_masks(tn, Hb, Wb,
                                                                                     Lb)
    xpre = model + off
    return xpre, Hb, Wb, Lb, T0, off
