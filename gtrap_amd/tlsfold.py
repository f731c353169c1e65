"""Periodic trapezoid least squares on the GPU (fold extension of gtrap/tls.py).

For each curve and each trial period P the light curve is folded, every epoch
``e < ceil(P/dtep)`` at ``t0 = e*dtep`` and every trapezoid (W, L) of the grid
is scored with the statistic of gtrap/tls.py:44-79 (``get_chi2Hmax`` with
``divx=True``), S/N = -H/sqrt(res/(n-3)) (tls.py:100), and the best trial of
each period is kept (first maximum in (W, L, epoch) order, NaN skipped, as
``tlsfit``'s nanargmax, tls.py:82-110).  The exact fold rule is
``u = fmod(t, P)``, ``d = u - t0`` wrapped by +-P into [-P/2, P/2);
``L = m * 0.45 * W / nl`` (the reference's normal-mode L grid,
include/gtls_simple.h:69,167).  Samples with a negative time (masked) or a
non-finite flux are skipped.  A curve longer than the batch's ``nmax`` is not
searched (S/N NaN, indices -2); :func:`pack` sets ``nmax`` to the longest curve.

The kernel is ``tlsfold_kernel`` (gtrap_amd/csrc/tlsfold.hip); this module
packs a ragged batch of curves, owns the workspace and returns host results.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native

DEFAULT_BLOCKS = 512


@dataclass
class FoldBatch:
    """A ragged batch resident on the device (flat t, x; per-curve offsets)."""
    t: object
    x: object
    off: object
    length: object
    periods: object
    poff: object
    ncurves: int
    nptot: int
    nmax: int
    poff_host: np.ndarray


def pack(curves, periods_list, device="cuda") -> FoldBatch:
    """Upload ``curves`` = [(t, x), ...] and per-curve period grids."""
    import torch
    lens = np.array([len(t) for t, _ in curves], np.int32)
    off = np.zeros(len(curves), np.int64)
    off[1:] = np.cumsum(lens[:-1])
    t = np.concatenate([np.asarray(t, np.float32) for t, _ in curves]) if curves else \
        np.zeros(0, np.float32)
    x = np.concatenate([np.asarray(x, np.float32) for _, x in curves]) if curves else \
        np.zeros(0, np.float32)
    nps = np.array([len(p) for p in periods_list], np.int64)
    poff = np.zeros(len(curves) + 1, np.int64)
    poff[1:] = np.cumsum(nps)
    per = np.concatenate([np.asarray(p, np.float64) for p in periods_list])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    return FoldBatch(dev(t), dev(x), dev(off), dev(lens), dev(per), dev(poff), len(curves),
                     int(poff[-1]), int(lens.max()) if len(lens) else 1, poff)


class Workspace:
    """Grow-only device workspace for tlsfold (per persistent workgroup)."""

    def __init__(self):
        self.buf = None

    def get(self, nmax: int, nblocks: int):
        import torch
        need = int(_native.load().gtrap_tlsfold_ws_bytes(int(nmax), int(nblocks)))
        if self.buf is None or self.buf.numel() < need:
            self.buf = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self.buf, need


_WS: dict = {}


def _workspace_for(stream) -> Workspace:
    """Default workspace of a stream: the work ticket and the per-workgroup
    slots live in it, so launches on different streams must not share one."""
    key = _native.stream_handle(stream)
    if key not in _WS:
        _WS[key] = Workspace()
    return _WS[key]


STATS = {"snr": 0, "dchi2": 1}


def tlsfold_device(batch: FoldBatch, widths, nl: int, dtep: float, out=None,
                   nblocks: int = DEFAULT_BLOCKS, workspace: Workspace | None = None,
                   stream=None, stat: str = "snr"):
    """Launch on a packed batch; returns device tensors (sn, h, w, l, e).
    ``stat``: "snr" (tls.py's S/N) or "dchi2" (the detection statistic, the
    fit's chi^2 improvement over the residual variance; reported in "sn")."""
    import torch
    from . import autotune
    _native.require_gpu()
    autotune.apply_tiles()
    ws = workspace or _workspace_for(stream)
    nb = max(1, min(int(nblocks), max(1, batch.nptot)))
    buf, need = ws.get(batch.nmax, nb)
    wdev = widths if isinstance(widths, torch.Tensor) else \
        torch.from_numpy(np.ascontiguousarray(widths, np.float64)).cuda()
    if out is None:
        out = dict(sn=torch.empty(batch.nptot, dtype=torch.float64, device="cuda"),
                   h=torch.empty(batch.nptot, dtype=torch.float64, device="cuda"),
                   w=torch.empty(batch.nptot, dtype=torch.int32, device="cuda"),
                   l=torch.empty(batch.nptot, dtype=torch.int32, device="cuda"),
                   e=torch.empty(batch.nptot, dtype=torch.int32, device="cuda"))
    p = _native.ptr
    _native.call("gtrap_tlsfold", p(out["sn"]), p(out["h"]), p(out["w"]), p(out["l"]),
                 p(out["e"]), p(batch.t), p(batch.x), p(batch.off), p(batch.length),
                 int(batch.nmax), p(batch.periods), p(batch.poff), int(batch.ncurves),
                 int(batch.nptot), p(wdev), int(wdev.numel()), int(nl), float(dtep),
                 STATS[stat], p(buf), int(need), nb, _native.stream_handle(stream))
    return out


def tlsfold(curves, periods_list, widths, nl: int = 10, dtep: float = 0.0204,
            nblocks: int = DEFAULT_BLOCKS, stat: str = "snr"):
    """Host convenience: returns a list of dicts (sn, h, w, l, e) per curve."""
    import torch
    batch = pack(curves, periods_list)
    out = tlsfold_device(batch, widths, nl, dtep, nblocks=nblocks, stat=stat)
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in out.items()}
    res = []
    for c in range(batch.ncurves):
        a, b = batch.poff_host[c], batch.poff_host[c + 1]
        res.append({k: v[a:b] for k, v in host.items()})
    return res


def period_grid(baseline: float, pmin: float = 0.5, nmin: int = 500, nmax: int = 4000,
                nlo: int = 2048, nhi: int = 70000, n: int | None = None):
    """Trial periods scaled to the baseline: geomspace(pmin, baseline/3, m) with m
    growing log-linearly from nmin (n = nlo samples) to nmax (n = nhi)."""
    if n is None:
        n = nlo
    f = np.clip((np.log(n) - np.log(nlo)) / (np.log(nhi) - np.log(nlo)), 0.0, 1.0)
    m = int(round(np.exp(np.log(nmin) + f * (np.log(nmax) - np.log(nmin)))))
    return np.geomspace(pmin, max(baseline / 3.0, pmin * 1.0001), m)
