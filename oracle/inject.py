"""Injection and random-number oracle (test infrastructure only).

* :func:`philox4x32` -- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
  random numbers: as easy as 1, 2, 3", SC 2011), the counter-based generator
  behind ``gtrap_gamma``: 10 rounds of two 32x32->64 multiplies with the
  published multipliers and Weyl key increments.  Pinned by the Random123
  known-answer vectors in tests/test_oracle_inject.py.
* :func:`uniform01`, :func:`normals` -- the mapping of Philox words to
  uniforms in (0, 1] and Box-Muller normals used by the kernel.
* :func:`trapezoid_profile` -- the injected dip: the model of
  gtrap/tls.py:2-15 (gen_trapzoid) shifted to 0 outside the transit, so
  depth H on the flat bottom [-W/4+L/2, W/4-L/2], linear ingress/egress,
  folded on the period around each transit centre.  Pinned against
  gen_trapzoid's formula restated directly in the test.

The gamma sampler (Marsaglia and Tsang 2000, with the u^(1/a) boost for
a < 1, as gtrap/ggenmock.py:18-79) has no bit-level reference (curand's
streams are not reproducible here); the GPU test checks its distribution
with a Kolmogorov-Smirnov test against scipy.stats.gamma ("parity
unpinned" at the bit level).
"""
from __future__ import annotations

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32(ctr, key, rounds=10):
    """ctr: 4 uint32, key: 2 uint32 -> 4 uint32 (Python ints)."""
    c0, c1, c2, c3 = (int(v) & MASK for v in ctr)
    k0, k1 = (int(v) & MASK for v in key)
    for r in range(rounds):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> 32, p0 & MASK
        hi1, lo1 = p1 >> 32, p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
        if r < rounds - 1:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def uniform01(word):
    """(word + 1) * 2^-32 in float64: uniform on (0, 1]."""
    return (float(word) + 1.0) * 2.0 ** -32


def trapezoid_profile(t, t0, period, depth, w, ell):
    """Dip of the tls.py trapezoid (total span W/2 + L, flat bottom W/2 - L)
    at every t, folded on ``period`` (None: single event at t0)."""
    t = np.asarray(t, np.float64)
    d = t - t0 if period is None else (t - t0 + 0.5 * period) % period - 0.5 * period
    a = np.abs(d)
    q4, h = w / 4.0, ell / 2.0
    prof = np.zeros_like(d)
    flat = a < q4 - h
    ramp = (a >= q4 - h) & (a < q4 + h)
    prof[flat] = depth
    prof[ramp] = depth * (q4 + h - a[ramp]) / ell
    return prof


def gamma(count, a, b, seed, max_try=10000):
    """CPU restatement of the gtrap_gamma sampler (same Philox words, same
    Marsaglia-Tsang steps, float64, rounded to float32)."""
    import math
    a_eff = a if a > 1.0 else a + 1.0
    d = a_eff - 1.0 / 3.0
    c = 1.0 / math.sqrt(9.0 * d)
    k0, k1 = seed & MASK, (seed >> 32) & MASK
    out = np.full(count, np.nan, np.float32)
    for i in range(count):
        for k in range(max_try):
            w = philox4x32((i & MASK, (i >> 32) & MASK, k, 0), (k0, k1))
            y = math.sqrt(-2.0 * math.log(uniform01(w[0]))) * math.cos(
                6.283185307179586 * uniform01(w[1]))
            v = 1.0 + c * y
            if v <= 0.0:
                continue
            v = v * v * v
            u = uniform01(w[2])
            if u < 1.0 - 0.0331 * (y * y) * (y * y) or \
                    math.log(u) < 0.5 * y * y + d * (1.0 - v + math.log(v)):
                g = d * v
                if a_eff > a:
                    g *= uniform01(w[3]) ** (1.0 / a)
                out[i] = np.float32(g / b)
                break
    return out


def inject(lc, tu, t0, period, depth, width, ell, mode="subtract"):
    """The injected image (interleaved (n, nq) float32), per-curve parameter
    arrays; period <= 0 is a single event; tu < 0 samples unchanged."""
    lc = np.array(lc, np.float32)
    tu = np.asarray(tu, np.float32)
    for q in range(lc.shape[1]):
        ok = tu[:, q] >= 0
        P = float(period[q])
        prof = trapezoid_profile(tu[ok, q].astype(np.float64), float(t0[q]),
                                 P if P > 0 else None, float(depth[q]), float(width[q]),
                                 float(ell[q]))
        x = lc[ok, q].astype(np.float64)
        lc[ok, q] = (x * (1.0 - prof) if mode == "multiply" else x - prof).astype(np.float32)
    return lc
