"""Single-transit-epoch trapezoid least squares: drop-in for gtrap/gtls.py.

``gtls_module(mode="normal"|"absolute").get_function("gtls")`` returns a
callable with the argument order of include/gtls_simple.h:1::

    k(dev_tlssn, dev_tlsw, dev_tlst0, dev_tlsl, dev_tlshmax, dev_lc, dev_tu,
      np.int32(nt), np.int32(nl), np.int32(nsc), np.float32(t0min),
      np.float32(dt), np.float32(wmax), np.float32(dw), np.float32(deltat),
      block=(nw, 1, 1), grid=(nt, nq), shared=S)

``block[0]`` = nw, the number of widths W = wmax - dw*iw (a power of two,
include/gtls_simple.h:383-390); ``grid = (nt, nq)``.  t0min and dt are
accepted and unused, as in the reference.  The kernels are
``gtls_fast_kernel`` (default: float64 statistic from tile prefix sums,
O(log nsc) per trial, the reference's float32 region decisions) and
``gtls_kernel`` (``arith="parity"``: the reference's float32 sequential scans,
bit for bit), both in gtrap_amd/csrc/gtls.hip.
"""
from __future__ import annotations

from . import _native
from ._module import Module

_MODE = {"normal": 0, "absolute": 1}


_ARITH = {"fast": 0, "parity": 2}


def gtls_device(tlssn, tlsw, tlst0, tlsl, tlshmax, lc, tu, nt, nq, nw, nl, nsc, wmax, dw,
                deltat, mode="normal", t0min=0.0, dt=0.0, stream=None, arith="fast") -> None:
    """arith="fast": float64 statistic, O(log nsc) per trial from prefix sums,
    the reference's float32 region decisions; arith="parity": the reference's
    float32 sequential scans, bit for bit."""
    _native.require_gpu()
    p = _native.ptr
    _native.call("gtrap_gtls", p(tlssn), p(tlsw), p(tlst0), p(tlsl), p(tlshmax), p(lc), p(tu),
                 int(nt), int(nq), int(nw), int(nl), int(nsc), float(t0min), float(dt),
                 float(wmax), float(dw), float(deltat), _MODE[mode] | _ARITH[arith],
                 _native.stream_handle(stream))


def gtls_module(mode: str = "normal", arith: str = "fast") -> Module:
    """Module with the ``gtls`` kernel (gtrap/gtls.py:11-27).  ``arith``
    selects the float64 fast evaluation (default) or the reference's float32
    scans (``"parity"``)."""
    if mode not in _MODE:
        mode = "normal"  # the reference treats any other string as normal mode
    if arith not in _ARITH:
        raise ValueError(f"arith must be one of {sorted(_ARITH)}")

    def kernel(dev_tlssn, dev_tlsw, dev_tlst0, dev_tlsl, dev_tlshmax, dev_lc, dev_tu, nt, nl,
               nsc, t0min, dt, wmax, dw, deltat, *, block, grid, shared, stream):
        if int(grid[0]) != int(nt):
            raise ValueError(f"gtls: grid[0]={grid[0]} must equal nt={int(nt)}")
        gtls_device(dev_tlssn, dev_tlsw, dev_tlst0, dev_tlsl, dev_tlshmax, dev_lc, dev_tu, nt,
                    grid[1], block[0], nl, nsc, wmax, dw, deltat, mode=mode, t0min=t0min, dt=dt,
                    stream=stream, arith=arith)

    return Module({"gtls": kernel})


def gtls_params(tu0, wmin, wmax, nw, nl):
    """Caller-side parameters of kepler/gtls_mockkepler.py:235-256, 295:
    deltat from curve 0's valid range, dw = (wmax-wmin)/(nw-1),
    nsc = int(wmax/deltat + 3), shared = (2*nsc + nw + 2)*4."""
    import numpy as np
    tt = np.asarray(tu0, np.float32)
    idx = np.arange(len(tt))
    m = tt >= 0.0
    deltat = (tt[m][-1] - tt[m][0]) / (idx[m][-1] - idx[m][0])
    dw = (wmax - wmin) / (nw - 1)
    nsc = int(wmax / deltat + 3.0)
    return dict(deltat=np.float32(deltat), dw=np.float32(dw), nsc=nsc, nl=nl, nw=nw,
                wmax=np.float32(wmax), shared=(2 * nsc + nw + 2) * 4)
