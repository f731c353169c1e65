"""Box least squares on a common time array: drop-in for gtrap/geebls_simple.py.

``gbls_module().get_function("geebls")(dev_blsp, dev_lc, dev_tu, n, nf, nb, kma,
kmi, kkmi, fmin, df, block=(nthread,1,1), grid=(nf, nq), shared=S)`` with the
argument order of gtrap/geebls_simple.py:17: 1-D ``tu[n]`` shared by all
curves, no mask, no offset (the caller pre-centres lc), N = n, SR only.
"""
from __future__ import annotations

from . import _native
from ._module import Module


def geebls_simple_device(blsp, lc, tu, n, nq, nf, nb, kma, kmi, kkmi, fmin, df, nthread=512,
                         stream=None, workspace=None) -> None:
    """``workspace`` as in :func:`gtrap_amd.geebls.geebls_device` (default: the
    per-stream one; ``False``: the interleaved image in place)."""
    from .geebls import workspace_for
    _native.require_gpu()
    p = _native.ptr
    args = (p(blsp), p(lc), p(tu), int(n), int(nq), int(nf), int(nb), int(kma), int(kmi),
            int(kkmi), float(fmin), float(df), int(nthread))
    if workspace is False:
        _native.call("gtrap_geebls_simple", *args, _native.stream_handle(stream))
        return
    buf, need = (workspace or workspace_for(stream)).get(n, nq, True)
    _native.call("gtrap_geebls_simple_ws", *args, p(buf), int(need),
                 _native.stream_handle(stream))


def _geebls_kernel(dev_blsp, dev_lc, dev_tu, n, nf, nb, kma, kmi, kkmi, fmin, df, *, block,
                   grid, shared, stream):
    if int(grid[0]) != int(nf):
        raise ValueError(f"geebls: grid[0]={grid[0]} must equal nf={int(nf)}")
    geebls_simple_device(dev_blsp, dev_lc, dev_tu, n, grid[1], nf, nb, kma, kmi, kkmi, fmin, df,
                         nthread=block[0], stream=stream)


def gbls_module() -> Module:
    """Module with the simple ``geebls`` kernel (gtrap/geebls_simple.py:10-179)."""
    return Module({"geebls": _geebls_kernel})
