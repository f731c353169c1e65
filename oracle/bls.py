"""BLS oracles (include/geebls.h:1-167, gtrap/geebls_simple.py:17-175).

* :func:`geebls` -- C restatement of the reference kernel (float32 arithmetic
  in the reference's order; bin sums exact in fixed point, rounded once).
* :func:`geebls_brute64` -- float64 brute force over every (start, width)
  box from the folded samples directly (tiny inputs), for pinning.

Pinned by tests/test_oracle_bls.py: brute force agreement, closed form on a
noiseless box, invariants (sign flip, tail samples ignored).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._clib import fptr, lib


def geebls(lc, tu, offset, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread=512, simple=False,
           qs=None, fs=None):
    """Returns (blsp, phase1, phase2) with shape (nq, nf) (curve-major)."""
    lc = np.ascontiguousarray(lc, np.float32)
    n, nq = lc.shape
    tu = np.ascontiguousarray(tu, np.float32)
    off = np.ascontiguousarray(np.zeros(nq) if offset is None else offset, np.float32)
    nt = np.ascontiguousarray(np.full(nq, n) if ntrue is None else ntrue, np.uint32)
    blsp = np.zeros(nf * nq, np.float32)
    p1 = np.full(nf * nq, 0xFFFFFFFF, np.uint32)
    p2 = np.full(nf * nq, 0xFFFFFFFF, np.uint32)
    f = lib().orc_geebls
    f.restype = None
    u = C.c_uint32
    f.argtypes = [C.c_void_p] * 7 + [u] * 7 + [C.c_float, C.c_float, u, C.c_int] + [u] * 4
    q0, q1 = (0, nq) if qs is None else qs
    f0, f1 = (0, nf) if fs is None else fs
    f(fptr(blsp), fptr(p1), fptr(p2), fptr(lc), fptr(tu), fptr(off), fptr(nt), n, nq, nf, nb,
      kma, kmi, kkmi, np.float32(fmin), np.float32(df), nthread, 1 if simple else 0, q0, q1,
      f0, f1)
    return blsp.reshape(nq, nf), p1.reshape(nq, nf), p2.reshape(nq, nf)


def scale_exp(maxabs: float, n: int) -> int:
    f = lib().orc_bls_scale_exp
    f.restype = C.c_int
    f.argtypes = [C.c_float, C.c_uint32]
    return f(np.float32(maxabs), n)


def geebls_brute64(lc, tu, offset, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread=512,
                   simple=False):
    """float64 brute force over every (start bin, width) box (tiny inputs).

    Folding and binning use the reference's float32 rule (so bin membership is
    the same); box sums are exact float64 sums over the folded samples, built
    independently of the C restatement (np.bincount + cumulative sums).
    f0 = fmin + df*jf is formed as float64(df)*jf + fmin rounded to float32
    (equal to fmaf except at rounding ties).
    Returns (sr[nq, nf], arg[nq, nf, 2]) with arg = (start, end) bins.
    """
    lc = np.asarray(lc, np.float32)
    n, nq = lc.shape
    tu = np.asarray(tu, np.float32)
    out = np.zeros((nq, nf))
    arg = np.full((nq, nf, 2), -1, np.int64)
    mcount = int(np.float32(n) / np.float32(nthread))
    nused = min(mcount * nthread, n)
    kmi_e, kkmi_e = max(kmi, 1), max(kkmi, 5)
    for q in range(nq):
        t = (tu if simple else tu[:, q])[:nused]
        off = np.float32(0.0) if simple else np.float32(offset[q])
        v = (lc[:nused, q] - off).astype(np.float64)
        N = float(n if simple else ntrue[q])
        ok = np.ones(nused, bool) if simple else (t >= 0)
        for jf in range(nf):
            f0 = np.float32(np.float64(np.float32(df)) * jf + np.float64(np.float32(fmin)))
            ph = (t * f0).astype(np.float32)
            ph = (ph - np.trunc(ph)).astype(np.float32)
            fj = (np.float32(nb) * ph).astype(np.float32)
            j = np.where(fj > 0, fj, 0).astype(np.int64)
            sel = ok & (j < nb)
            y = np.bincount(j[sel], weights=v[sel], minlength=nb)
            c = np.bincount(j[sel], minlength=nb)
            cs = np.concatenate([[0.0], np.cumsum(np.concatenate([y, y[:kma]]))])
            cc = np.concatenate([[0], np.cumsum(np.concatenate([c, c[:kma]]))])
            best, bi, bj = 0.0, -1, -1
            for i in range(nb):
                for k in range(kmi_e + 1, kma + 1):
                    jj = i + k - 1
                    kk = cc[jj + 1] - cc[i]
                    if kk > kkmi_e:
                        s = cs[jj + 1] - cs[i]
                        pw = s * s / (kk * (N - kk))
                        if pw > best:
                            best, bi, bj = pw, i, jj
            out[q, jf] = np.sqrt(best)
            arg[q, jf] = (bi, bj)
    return out, arg
