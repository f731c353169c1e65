"""Box least squares with mask, offset and non-common time: drop-in for gtrap/geebls.py.

``gbls_module().get_function("geebls")`` returns a callable with the argument
order of include/geebls.h:1::

    k(dev_blsp, dev_phase1, dev_phase2, dev_lc, dev_tu, dev_offset, dev_ntrue,
      np.uint32(n), np.uint32(nf), np.uint32(nb), np.uint32(kma), np.uint32(kmi),
      np.uint32(kkmi), np.float32(fmin), np.float32(df),
      block=(nthread, 1, 1), grid=(nf, nq), shared=S)

``block[0]`` is the reference's thread count and keeps its meaning (samples
``i >= floor(n/nthread)*nthread`` are not binned, include/geebls.h:57);
``grid = (nf, nq)``; ``shared`` is ignored.  The kernel itself is
``bls_kernel`` in gtrap_amd/csrc/bls.hip.
"""
from __future__ import annotations

from . import _native
from ._module import Module


class Workspace:
    """Grow-only device workspace of the curve-major BLS path
    (``gtrap_geebls_ws_bytes``: the image copied curve-major plus one scale
    per curve)."""

    def __init__(self):
        self.buf = None

    def get(self, n: int, nq: int, simple: bool):
        import torch
        need = int(_native.load().gtrap_geebls_ws_bytes(int(n), int(nq), int(bool(simple))))
        if self.buf is None or self.buf.numel() < need:
            self.buf = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self.buf, need


_WS: dict = {}


def workspace_for(stream=None) -> Workspace:
    """This module's default workspace for a stream (calls on one stream are
    ordered, so they can share it)."""
    import torch
    key = (torch.cuda.current_device(), _native.stream_handle(stream))
    if key not in _WS:
        _WS[key] = Workspace()
    return _WS[key]


def geebls_device(blsp, phase1, phase2, lc, tu, offset, ntrue, n, nq, nf, nb, kma, kmi, kkmi,
                  fmin, df, nthread=512, stream=None, workspace=None) -> None:
    """Launch BLS on device buffers; interleaved (n, nq) lc/tu, outputs at jf + nf*q.

    ``workspace`` (default: this module's per-stream one) selects the
    curve-major path (``gtrap_geebls_ws``); ``workspace=False`` reads the
    interleaved image in place (``gtrap_geebls``, no device memory beyond the
    arguments).  Both give the same bits."""
    _native.require_gpu()
    p = _native.ptr
    args = (p(blsp), p(phase1), p(phase2), p(lc), p(tu), p(offset), p(ntrue), int(n), int(nq),
            int(nf), int(nb), int(kma), int(kmi), int(kkmi), float(fmin), float(df),
            int(nthread))
    if workspace is False:
        _native.call("gtrap_geebls", *args, _native.stream_handle(stream))
        return
    buf, need = (workspace or workspace_for(stream)).get(n, nq, False)
    _native.call("gtrap_geebls_ws", *args, p(buf), int(need), _native.stream_handle(stream))


def geebls_f64_device(sr, phase1, phase2, lc, tu, t64, offset, ntrue, n, nq, nf, nb, kma, kmi,
                      kkmi, fmin, df, nthread=512, stream=None) -> None:
    """Precision mode of geebls: the same search with a float64 fold
    (f = fmin + df*jf, ph = frac(t*f) from float64 time ``t64``, interleaved
    like ``tu``) and float64 bin sums, running sums and statistic; ``sr`` is
    float64 at jf + nf*q.  Samples are valid where tu >= 0.  Over long
    baselines the reference's float32 fold (include/geebls.h:68-70) moves
    samples across bins; this mode does not."""
    _native.require_gpu()
    p = _native.ptr
    _native.call("gtrap_geebls_prec", 14, p(sr), p(phase1), p(phase2), p(lc), None, None, p(tu),
                 p(t64), p(offset), p(ntrue), int(n), int(nq), int(nf), int(nb), int(kma),
                 int(kmi), int(kkmi), float(fmin), float(df), int(nthread),
                 _native.stream_handle(stream))


def _geebls_kernel(dev_blsp, dev_phase1, dev_phase2, dev_lc, dev_tu, dev_offset, dev_ntrue,
                   n, nf, nb, kma, kmi, kkmi, fmin, df, *, block, grid, shared, stream):
    if int(grid[0]) != int(nf):
        raise ValueError(f"geebls: grid[0]={grid[0]} must equal nf={int(nf)}")
    geebls_device(dev_blsp, dev_phase1, dev_phase2, dev_lc, dev_tu, dev_offset, dev_ntrue,
                  n, grid[1], nf, nb, kma, kmi, kkmi, fmin, df, nthread=block[0], stream=stream)


def gbls_module() -> Module:
    """Module with the ``geebls`` kernel (gtrap/geebls.py:10-22)."""
    return Module({"geebls": _geebls_kernel})


def bls_params(n, nb, qmi, qma):
    """Caller-side parameters of the reference notebooks (A10):
    kma = int(qma*nb)+1, kmi = int(qmi*nb), kkmi = int(n*qmi) and the LDS size
    4*(nb+kma)*2 + 4*nb the reference passes as ``shared``."""
    kma = int(qma * nb) + 1
    kmi = int(qmi * nb)
    kkmi = int(n * qmi)
    return kma, kmi, kkmi, int(4 * (nb + kma) * 2 + 4 * nb)
