"""A small PyCUDA-style driver shim over torch-ROCm device memory.

The reference's callers use ``pycuda.driver`` as ``cuda``: ``cuda.mem_alloc``,
``cuda.memcpy_htod`` and ``cuda.memcpy_dtoh`` (e.g. gtrap/gfilter.py:125-135,
kepler/gtls_mockkepler.py:269-318).  This module offers the same three calls so
those scripts run with ``import gtrap_amd.driver as cuda``.  Memory is a torch
uint8 tensor; torch's caching allocator owns it and frees it on GC, like
PyCUDA's DeviceAllocation.
"""
from __future__ import annotations

import numpy as np


class DeviceAllocation:
    """Device buffer of ``nbytes`` bytes; ``int(a)`` is its address."""

    __slots__ = ("tensor",)

    def __init__(self, nbytes: int, device: str = "cuda"):
        import torch
        self.tensor = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)

    @classmethod
    def wrap(cls, tensor) -> "DeviceAllocation":
        a = cls.__new__(cls)
        a.tensor = tensor.view(-1).view(__import__("torch").uint8)
        return a

    @property
    def nbytes(self) -> int:
        return self.tensor.numel()

    def __int__(self) -> int:
        return self.tensor.data_ptr()

    def data_ptr(self) -> int:
        return self.tensor.data_ptr()

    @property
    def is_cuda(self) -> bool:
        return True

    def view(self, dtype, count: int | None = None):
        """A typed torch view of the buffer (numpy or torch dtype)."""
        import torch
        tdt = _torch_dtype(dtype)
        esz = torch.empty((), dtype=tdt).element_size()
        cnt = self.nbytes // esz if count is None else int(count)
        return self.tensor[: cnt * esz].view(tdt)


def _torch_dtype(dtype):
    import torch
    if isinstance(dtype, torch.dtype):
        return dtype
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
            np.dtype(np.uint32): torch.uint32, np.dtype(np.int32): torch.int32,
            np.dtype(np.int64): torch.int64, np.dtype(np.uint8): torch.uint8,
            np.dtype(np.float16): torch.float16}[np.dtype(dtype)]


def mem_alloc(nbytes: int) -> DeviceAllocation:
    return DeviceAllocation(nbytes)


def _as_bytes_tensor(dev):
    import torch
    if isinstance(dev, DeviceAllocation):
        return dev.tensor
    if isinstance(dev, torch.Tensor):
        return dev.view(-1).view(torch.uint8)
    raise TypeError(f"not a device buffer: {type(dev).__name__}")


def memcpy_htod(dev, host) -> None:
    """Copy a host array into a device buffer (synchronous, like PyCUDA)."""
    import torch
    h = np.ascontiguousarray(host)
    dst = _as_bytes_tensor(dev)
    if h.nbytes > dst.numel():
        raise ValueError(f"memcpy_htod: {h.nbytes} bytes into a {dst.numel()}-byte buffer")
    src = torch.from_numpy(h.view(np.uint8).reshape(-1))
    dst[: h.nbytes].copy_(src)
    torch.cuda.current_stream().synchronize()


def memcpy_dtoh(host: np.ndarray, dev) -> None:
    """Copy a device buffer into a host array (synchronises the stream)."""
    import torch
    if not host.flags["C_CONTIGUOUS"]:
        raise ValueError("memcpy_dtoh: host array must be C-contiguous")
    src = _as_bytes_tensor(dev)
    nb = host.nbytes
    if nb > src.numel():
        raise ValueError(f"memcpy_dtoh: {nb} bytes from a {src.numel()}-byte buffer")
    out = src[:nb].cpu().numpy()
    host.view(np.uint8).reshape(-1)[:] = out


def to_device(host, dtype=None):
    """Upload a host array as a new torch device tensor (helper)."""
    import torch
    h = np.ascontiguousarray(host if dtype is None else np.asarray(host, dtype=dtype))
    return torch.from_numpy(h).to("cuda")
