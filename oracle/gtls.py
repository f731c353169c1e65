"""gtls oracle (include/gtls_simple.h, include/gtls_absolute.h).

:func:`gtls` wraps the float32 line-faithful C restatement ``orc_gtls``.
Pinned by tests/test_oracle_gtls.py: agreement with the float64 tls.py
statistic where the reference's quirks do not apply, the region-A index
quirk (A.1) being observable, fill-factor and masking rules.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._clib import fptr, lib


def gtls(lc, tu, nw, nl, nsc, wmax, dw, deltat, mode="normal", ks=None):
    lc = np.ascontiguousarray(lc, np.float32)
    tu = np.ascontiguousarray(tu, np.float32)
    nt, nq = lc.shape
    outs = [np.zeros(nt * nq, np.float32) for _ in range(5)]
    f = lib().orc_gtls
    f.restype = None
    f.argtypes = [C.c_void_p] * 7 + [C.c_int] * 5 + [C.c_float] * 3 + [C.c_int] * 3
    k0, k1 = (0, nt) if ks is None else ks
    f(*(fptr(o) for o in outs), fptr(lc), fptr(tu), nt, nq, nw, nl, nsc, np.float32(wmax),
      np.float32(dw), np.float32(deltat), 1 if mode == "absolute" else 0, k0, k1)
    names = ("tlssn", "tlsw", "tlst0", "tlsl", "tlshmax")
    return {k: o.reshape(nt, nq) for k, o in zip(names, outs)}


def gtls64(lc, tu, nw, nl, nsc, wmax, dw, deltat, mode="normal", ks=None):
    """gtls with the reference's float32 region decisions and float64
    arithmetic (``orc_gtls64``); also returns the float64 best S/N as
    ``sn64``."""
    lc = np.ascontiguousarray(lc, np.float32)
    tu = np.ascontiguousarray(tu, np.float32)
    nt, nq = lc.shape
    outs = [np.zeros(nt * nq, np.float32) for _ in range(5)]
    sn64 = np.zeros(nt * nq, np.float64)
    f = lib().orc_gtls64
    f.restype = None
    f.argtypes = [C.c_void_p] * 8 + [C.c_int] * 5 + [C.c_float] * 3 + [C.c_int] * 3
    k0, k1 = (0, nt) if ks is None else ks
    f(*(fptr(o) for o in outs), fptr(sn64), fptr(lc), fptr(tu), nt, nq, nw, nl, nsc,
      np.float32(wmax), np.float32(dw), np.float32(deltat), 1 if mode == "absolute" else 0, k0, k1)
    names = ("tlssn", "tlsw", "tlst0", "tlsl", "tlshmax")
    out = {k: o.reshape(nt, nq) for k, o in zip(names, outs)}
    out["sn64"] = sn64.reshape(nt, nq)
    return out
