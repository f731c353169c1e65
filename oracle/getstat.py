"""BLS statistics oracle (gtrap/getstat.py:14-123), float64 restatement.

Per curve over the first floor(nf/patch)*patch frequencies: the lowest index
of the maximum SR (getstat.py:46-48,109-110), Pest = 1/(idx*df + fmin)
(:111), ave = sum/nf, var = sumsq/nf - ave^2 (:114-116),
SDE = (max - ave)/sqrt(var) (:118), phases at idx (:120-121).
Pinned by tests/test_oracle_getstat.py against a closed form and against the
reference's patch-then-combine structure restated separately.
"""
from __future__ import annotations

import numpy as np


def blsstat(blsp, phase1, phase2, nf, nq, df, fmin, patch=1024):
    blsp = np.asarray(blsp, np.float32).reshape(nq, nf)
    used = (nf // patch) * patch
    b = blsp[:, :used].astype(np.float64)
    idx = np.argmax(blsp[:, :used], axis=1)
    mx = blsp[np.arange(nq), idx].astype(np.float64)
    ave = b.sum(1) / nf
    var = (b * b).sum(1) / nf - ave * ave
    sde = (mx - ave) / np.sqrt(var)
    pest = 1.0 / (idx * df + fmin)
    p1 = None if phase1 is None else np.asarray(phase1).reshape(nq, nf)[np.arange(nq), idx]
    p2 = None if phase2 is None else np.asarray(phase2).reshape(nq, nf)[np.arange(nq), idx]
    return pest, sde, p1, p2, idx


def blsstat_patchwise(blsp, nf, nq, df, fmin, patch=1024):
    """The reference's two-stage structure: per-patch (max, sum, sumsq, first
    argmax) then combine on the host (gtrap/getstat.py:69-123), in float64."""
    blsp = np.asarray(blsp, np.float32).reshape(nq, nf)
    npatch = nf // patch
    pest, sde = [], []
    for q in range(nq):
        pm, ps, pss, pi = [], [], [], []
        for k in range(npatch):
            seg = blsp[q, k * patch:(k + 1) * patch].astype(np.float64)
            j = int(np.argmax(seg))
            pm.append(seg[j]); ps.append(seg.sum()); pss.append((seg * seg).sum())
            pi.append(k * patch + j)
        kp = int(np.argmax(pm))
        idx = pi[kp]
        ave = np.sum(ps) / nf
        var = np.sum(pss) / nf - ave * ave
        pest.append(1.0 / (idx * df + fmin))
        sde.append((pm[kp] - ave) / np.sqrt(var))
    return np.array(pest), np.array(sde)
