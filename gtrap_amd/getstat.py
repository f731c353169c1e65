"""BLS statistics: drop-in for gtrap/getstat.py.

``get_blsstat(dev_blsp, dev_phase1, dev_phase2, nf, nq, df, fmin, nthread=1024)``
returns ``(Pest, sde, phasebest1, phasebest2)`` host arrays like
gtrap/getstat.py:64-67.  The reduction runs on the device in float64
(``blsstats_kernel`` in gtrap_amd/csrc/stats.hip); only the four per-curve
results are copied back (the reference copies six per-patch arrays and
combines them on the host, getstat.py:91-123).

Per curve, over the first floor(nf/nthread)*nthread frequencies:
idx = lowest index of the max SR, Pest = 1/(idx*df + fmin),
ave = sum/nf, var = sumsq/nf - ave^2, SDE = (max - ave)/sqrt(var).
"""
from __future__ import annotations

import numpy as np

from . import _native


def blsstats_device(blsp, phase1, phase2, nf, nq, df, fmin, patch=1024, stream=None):
    """Device outputs: (pest f64, sde f64, phasebest1 u32, phasebest2 u32, maxval f32)."""
    import torch
    _native.require_gpu()
    nq = int(nq)
    pest = torch.empty(nq, dtype=torch.float64, device="cuda")
    sde = torch.empty(nq, dtype=torch.float64, device="cuda")
    pb1 = torch.empty(nq, dtype=torch.int32, device="cuda")
    pb2 = torch.empty(nq, dtype=torch.int32, device="cuda")
    mx = torch.empty(nq, dtype=torch.float32, device="cuda")
    p = _native.ptr
    _native.call("gtrap_blsstats", p(blsp), p(phase1), p(phase2), int(nf), nq, int(patch),
                 float(df), float(fmin), p(pest), p(sde), p(pb1), p(pb2), p(mx),
                 _native.stream_handle(stream))
    return pest, sde, pb1, pb2, mx


def get_blsstat(dev_blsp, dev_phase1, dev_phase2, nf, nq, df, fmin, nthread=1024):
    """gtrap/getstat.py:64-67 -> (Pest, sde, phasebest1, phasebest2)."""
    pest, sde, pb1, pb2, _ = blsstats_device(dev_blsp, dev_phase1, dev_phase2, nf, nq, df,
                                             fmin, patch=nthread)
    return (pest.cpu().numpy(), sde.cpu().numpy(),
            pb1.cpu().numpy().view(np.uint32), pb2.cpu().numpy().view(np.uint32))
