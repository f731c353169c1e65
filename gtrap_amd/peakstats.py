"""Post-search peak statistics on the GPU: the block after the single-transit
search in gtrap/gtls.py:244-275 and kepler/gtls_mockkepler.py:336-376, with
gtrap/detect_peaks.py's peak finder, for every curve of a batch at once.

``peakstats_device(tlssn, nt, nq, mpd=48, ntop=4, tlst0=None)`` takes the
interleaved (k*nq + q) device outputs of gtls and returns device tensors;
``get_peakstats`` returns a host dict per statistic (arrays of length nq) and
the top peak's T0, W, L, H when the gtls outputs are given.
"""
from __future__ import annotations

import numpy as np

from . import _native

NAMES = ("maxsn", "median", "std", "far", "diff", "pssn", "pinterval")


def peakstats_device(tlssn, nt, nq, mpd=48, ntop=4, tlst0=None, nblocks=256, stream=None):
    import torch
    _native.require_gpu()
    nt, nq = int(nt), int(nq)
    nb = max(1, min(int(nblocks), nq))
    need = int(_native.load().gtrap_peakstats_ws_bytes(nt, nb))
    ws = torch.empty(max(need, 8), dtype=torch.uint8, device="cuda")
    st = torch.empty((nq, 8), dtype=torch.float64, device="cuda")
    cnt = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
    top = torch.empty((nq, int(ntop)), dtype=torch.int32, device="cuda")
    p = _native.ptr
    _native.call("gtrap_peakstats", p(st), p(cnt), p(top), p(tlssn), p(tlst0), nt, nq, int(mpd),
                 int(ntop), p(ws), need, nb, _native.stream_handle(stream))
    return st, cnt, top


def get_peakstats(tlssn, nt, nq, mpd=48, ntop=4, tlst0=None, tlsw=None, tlsl=None,
                  tlshmax=None):
    st, cnt, top = peakstats_device(tlssn, nt, nq, mpd, ntop, tlst0)
    st, cnt, top = st.cpu().numpy(), cnt.cpu().numpy(), top.cpu().numpy()
    out = {k: st[:, i] for i, k in enumerate(NAMES)}
    out.update(npeak=cnt[:, 0], lent=cnt[:, 1], top=top)
    k0 = top[:, 0]
    for name, arr in (("T0", tlst0), ("W", tlsw), ("L", tlsl), ("H", tlshmax)):
        if arr is not None:
            a = arr.reshape(-1)
            idx = np.maximum(k0, 0).astype(np.int64) * nq + np.arange(nq)
            v = a[idx].cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)[idx]
            out[name] = np.where(k0 >= 0, v, np.nan)
    return out
