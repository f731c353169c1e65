"""float64 BLS yardstick for the precision study (test infrastructure only).

Restates the box search of include/geebls.h:1-167 in float64 with a float64
fold: f = fmin + df*jf, ph = t*f - floor(t*f), bin j = int(nb*ph) (0 for a
non-positive product), samples i < floor(n/nthread)*nthread with tu >= 0,
bins extended by their first kma entries, pow = s^2 / (kk (ntrue - kk)) for
k > max(kmi, 1) and kk > max(kkmi, 5), SR = sqrt(max pow).  Bin sums are
np.bincount sums in float64 and box sums are cumulative-sum differences, so
the result agrees with the GPU's exact-sum float64 variant to rounding
(~1e-13 relative), not bit for bit.

Pinned by tests/test_oracle_precision.py: equals oracle.bls.geebls_brute64
where the float64 fold bins like the float32 one (coarse grid, short time
span), the closed form of a noiseless box, and invariance under a sign flip.
"""
from __future__ import annotations

import numpy as np


def bls64_cells(lc, t64, tu, offset, ntrue, nb, kma, kmi, kkmi, fmin, df, nthread, cells):
    """SR for each (q, jf) in ``cells``; lc, tu (n, nq) float32, t64 (n,) or
    (n, nq) float64, offset float32[nq], ntrue[nq]."""
    lc = np.asarray(lc, np.float32)
    n, nq = lc.shape
    t64 = np.asarray(t64, np.float64)
    tu = np.asarray(tu, np.float32)
    mcount = int(np.float32(n) / np.float32(nthread))
    nused = min(mcount * nthread, n)
    kmi_e, kkmi_e = max(kmi, 1), max(kkmi, 5)
    out = []
    for q, jf in cells:
        t = (t64 if t64.ndim == 1 else t64[:, q])[:nused]
        ok = tu[:nused, q] >= 0
        v = (lc[:nused, q] - np.float32(offset[q])).astype(np.float64)
        N = float(ntrue[q])
        f0 = fmin + df * float(jf)
        ph = t * f0
        ph = ph - np.floor(ph)
        fj = float(nb) * ph
        j = np.where(fj > 0, fj, 0).astype(np.int64)
        sel = ok & (j < nb)
        y = np.bincount(j[sel], weights=v[sel], minlength=nb)
        c = np.bincount(j[sel], minlength=nb)
        cs = np.concatenate([[0.0], np.cumsum(np.concatenate([y, y[:kma]]))])
        cc = np.concatenate([[0], np.cumsum(np.concatenate([c, c[:kma]]))])
        best = 0.0
        i = np.arange(nb)
        for k in range(kmi_e + 1, kma + 1):
            kk = cc[i + k] - cc[i]
            s = cs[i + k] - cs[i]
            m = kk > kkmi_e
            if m.any():
                pw = s[m] ** 2 / (kk[m] * (N - kk[m]))
                best = max(best, float(pw.max()))
        out.append(np.sqrt(best))
    return np.array(out)
