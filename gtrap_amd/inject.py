"""Synthetic injection on the GPU: gamma deviates (gtrap/ggenmock.py:18-79)
and trapezoid transits on an interleaved (n, nq) image (the role of
gtrap/genmock.py:9-43 with the gtrap/tls.py:2-15 trapezoid instead of a
batman limb-darkened model)."""
from __future__ import annotations

import numpy as np

from . import _native


def gamma(count: int, a: float, b: float = 1.0, seed: int = 10, stream=None):
    """Device float32 tensor of ``count`` Gamma(a, rate b) deviates; the value
    at index i depends only on (seed, i).  Raises if a slot exhausted its
    proposals (it would be NaN)."""
    import torch
    _native.require_gpu()
    x = torch.empty(int(count), dtype=torch.float32, device="cuda")
    nfail = torch.empty(1, dtype=torch.int32, device="cuda")
    _native.call("gtrap_gamma", _native.ptr(x), int(count), float(a), float(b),
                 int(seed) & 0xFFFFFFFFFFFFFFFF, _native.ptr(nfail), _native.stream_handle(stream))
    if int(nfail.item()):
        raise RuntimeError(f"gamma: {int(nfail.item())} deviates exhausted their proposals")
    return x


def inject_trapezoid(lc, tu, t0, period, depth, width, ell, mode: str = "subtract",
                     stream=None) -> None:
    """In place on device lc (n, nq) float32 with tu (n, nq): per-curve
    parameter sequences (host or device, length nq); period <= 0 injects one
    event at t0.  mode "subtract" (flux - dip) or "multiply" (flux * (1 - dip))."""
    import torch
    _native.require_gpu()
    n, nq = int(lc.shape[0]), int(lc.shape[1])
    dev = lambda v: (v if isinstance(v, torch.Tensor) else
                     torch.from_numpy(np.ascontiguousarray(np.broadcast_to(
                         np.asarray(v, np.float64), (nq,))))).to("cuda", torch.float64)
    args = [dev(v) for v in (t0, period, depth, width, ell)]
    if mode not in ("subtract", "multiply"):
        raise ValueError("mode must be 'subtract' or 'multiply'")
    p = _native.ptr
    _native.call("gtrap_inject_trapezoid", p(lc), p(tu), n, nq, *(p(a) for a in args),
                 1 if mode == "multiply" else 0, _native.stream_handle(stream))
