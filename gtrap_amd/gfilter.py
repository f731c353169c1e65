"""Median-filter detrend: drop-in for gtrap/gfilter.py.

Same names and argument meaning as the reference module:

* ``gmf_module()`` -> module with kernel ``"gfilter"`` taking
  ``(dev_in, dev_out, uint32 i_dim, uint32 j_dim, uint32 r, block=(nbx, nby, 1),
  grid=(j_dim/nbx, i_dim/nby), shared=S)`` (gtrap/gfilter.py:13,131).
* ``get_offset(n, nq, dev_imgout, tu)`` -> ``(offset, imgout)`` (gtrap/gfilter.py:94-101).
* ``get_detrend(img, nby, nbx, r, isw=0, osw=0)`` (gtrap/gfilter.py:103-143).

The output is ``x - median`` (the reference kernel, gtrap/gfilter.py:76) by
default, or ``x / median`` with ``mode="div"``.  The median is the exact
zero-padded window median (identical to ``scipy.signal.medfilt``).
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _native
from ._module import Module
from .driver import DeviceAllocation, memcpy_dtoh, memcpy_htod

_MODES = {"sub": 0, "div": 1, 0: 0, 1: 1}


def _mode(mode) -> int:
    try:
        return _MODES[mode]
    except KeyError:
        raise ValueError(f"mode must be 'sub' or 'div', got {mode!r}") from None


def medfilt_device(dev_in, dev_out, n: int, nq: int, r: int, mode="sub",
                   rows_out: int | None = None, cols_out: int | None = None,
                   seg_len: int = 0, stream=None) -> None:
    """Launch the detrend on device buffers holding an interleaved (n, nq) image."""
    _native.require_gpu()
    rows = n if rows_out is None else min(int(rows_out), n)
    cols = nq if cols_out is None else min(int(cols_out), nq)
    _native.call("gtrap_medfilt", _native.ptr(dev_in), _native.ptr(dev_out), int(n), int(nq),
                 int(r), _mode(mode), rows, cols, int(seg_len), _native.stream_handle(stream))


def _gfilter_kernel(dev_in, dev_out, i_dim, j_dim, r, *, block, grid, shared, stream,
                    mode="sub", strict=False):
    """The ``gfilter`` kernel call of gtrap/gfilter.py:131.

    Rows ``i < grid[1]*block[1]`` and columns ``q < grid[0]*block[0]`` are
    written, as in the reference; ``i_dim`` (= n) bounds the zero padding.
    """
    nbx, nby = block[0], block[1]
    i_dim, j_dim, r = int(i_dim), int(j_dim), int(r)
    if strict:
        if nby < 2 * r:
            raise ValueError(f"gfilter: block[1]={nby} < 2r={2 * r} (reference halo undefined)")
        if i_dim % nby or j_dim % nbx:
            raise ValueError("gfilter: i_dim/j_dim not divisible by the block shape")
    medfilt_device(dev_in, dev_out, i_dim, j_dim, r, mode,
                   rows_out=grid[1] * nby, cols_out=grid[0] * nbx, stream=stream)


def gmf_module() -> Module:
    """Module with the ``gfilter`` kernel (gtrap/gfilter.py:8-89)."""
    return Module({"gfilter": _gfilter_kernel})


def maxpow(num):
    """Next power of two (gtrap/gfilter.py:91-92)."""
    return 2 ** int(np.ceil(np.log2(num)))


def get_offset(n, nq, dev_imgout, tu, *, device_only=False):
    """Masked per-curve mean over n samples (gtrap/gfilter.py:94-101).

    Returns ``(offset, imgout)`` like the reference: ``offset`` (float32, length
    nq) is sum over unmasked samples divided by ``n`` (not by the number of
    unmasked samples), computed on the device in float64; ``imgout`` is the
    (n, nq) host image with masked samples set to 0.  With ``device_only=True``
    returns ``(offset32_dev, offset64_dev, ntrue_dev)`` torch tensors instead and
    copies nothing to the host.
    """
    import torch
    _native.require_gpu()
    n, nq = int(n), int(nq)
    tu_dev = tu if isinstance(tu, torch.Tensor) and tu.is_cuda else \
        torch.from_numpy(np.ascontiguousarray(tu, dtype=np.float32).reshape(-1)).cuda()
    off32 = torch.empty(nq, dtype=torch.float32, device="cuda")
    off64 = torch.empty(nq, dtype=torch.float64, device="cuda")
    ntrue = torch.empty(nq, dtype=torch.int32, device="cuda")
    _native.call("gtrap_offset", _native.ptr(dev_imgout), _native.ptr(tu_dev), n, nq,
                 _native.ptr(off64), _native.ptr(off32), _native.ptr(ntrue),
                 _native.stream_handle())
    if device_only:
        return off32, off64, ntrue
    imgout = np.zeros(nq * n, dtype=np.float32)
    memcpy_dtoh(imgout, dev_imgout)
    tuh = tu_dev.cpu().numpy().reshape(-1)
    imgout[tuh < 0.0] = 0.0
    return off32.cpu().numpy(), imgout.reshape((n, nq))


def get_detrend(img, nby, nbx, r, isw=0, osw=0, *, mode="sub", strict=False, verbose=False):
    """Upload an (n, nq) image, run the detrend, return host array or device buffer.

    gtrap/gfilter.py:103-143: ``osw=0`` returns the (ydim, xdim) float32 host
    array, ``osw=1`` returns the device buffer.  ``isw`` must be 0 (the
    reference's other branch references an undefined buffer).  The reference
    prints warnings for a non-Fortran array and for shapes not divisible by
    the block; ``strict=True`` turns those into errors.
    """
    if isw != 0:
        raise ValueError("get_detrend: only isw=0 is defined (gtrap/gfilter.py:124-126)")
    img = np.asarray(img)
    if img.ndim != 2:
        raise ValueError("get_detrend: img must be a 2-D (n, nq) array")
    ydim, xdim = img.shape
    if not np.isfortran(img):
        msg = "gfilter assumes a Fortran-wise batch array. Probably, wrong results?"
        if strict:
            raise ValueError(msg)
        if verbose:
            warnings.warn(msg)
    if (xdim % nbx or ydim % nby) and strict:
        raise ValueError("get_detrend: image shape not divisible by the block shape")
    flat = np.ascontiguousarray(img.flatten().astype(np.float32))
    dev_in = DeviceAllocation(flat.nbytes)
    memcpy_htod(dev_in, flat)
    dev_out = DeviceAllocation(flat.nbytes)
    gmf_module().get_function("gfilter")(
        dev_in, dev_out, np.uint32(ydim), np.uint32(xdim), np.uint32(r),
        block=(int(nbx), int(nby), 1), grid=(int(xdim // nbx), int(ydim // nby)),
        mode=mode, strict=strict)
    if osw == 0:
        out1 = np.zeros(xdim * ydim, dtype=np.float32)
        memcpy_dtoh(out1, dev_out)
        return out1.reshape((ydim, xdim))
    return dev_out
