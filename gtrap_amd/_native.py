"""ctypes binding of libgtrap_hip.so (the C ABI declared in include/gtrap_hip.h).

Device buffers are torch tensors (or :class:`gtrap_amd.driver.DeviceAllocation`)
on a ROCm device; only their addresses cross the ABI.  There is no CPU fallback:
if the library is missing or no GPU is visible, every compute call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libgtrap_hip.so")

_vp, _u32, _i32, _i64, _f32, _f64, _int = (C.c_void_p, C.c_uint32, C.c_int32, C.c_int64,
                                           C.c_float, C.c_double, C.c_int)

# name -> (restype, argtypes); must match include/gtrap_hip.h
SIGNATURES: dict[str, tuple] = {
    "gtrap_last_error": (C.c_char_p, []),
    "gtrap_version": (C.c_char_p, []),
    "gtrap_medfilt": (_int, [_vp, _vp, _u32, _u32, _u32, _int, _u32, _u32, _u32, _vp]),
    "gtrap_offset": (_int, [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "gtrap_medfilt_ragged": (_int, [_vp] * 7 + [_i32, _u32, _int, _i32, _vp]),
    "gtrap_geebls": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32,
                            _u32, _u32, _f32, _f32, _u32, _vp]),
    "gtrap_geebls_simple": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32,
                                   _f32, _f32, _u32, _vp]),
    "gtrap_geebls_ws_bytes": (_i64, [_u32, _u32, _i32]),
    "gtrap_geebls_ws": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32,
                               _u32, _u32, _f32, _f32, _u32, _vp, _i64, _vp]),
    "gtrap_geebls_simple_ws": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32,
                                      _f32, _f32, _u32, _vp, _i64, _vp]),
    "gtrap_geebls_prec": (_int, [_int] + [_vp] * 10 + [_u32] * 7 + [_f64, _f64, _u32, _vp]),
    "gtrap_image_build": (_int, [_vp] * 11 + [_i32, _i32, _i32, _f32, _f32, _int, _vp, _vp]),
    "gtrap_gamma": (_int, [_vp, _i64, _f64, _f64, C.c_uint64, _vp, _vp]),
    "gtrap_inject_trapezoid": (_int, [_vp, _vp, _i32, _i32] + [_vp] * 5 + [_int, _vp]),
    "gtrap_peakstats_ws_bytes": (_i64, [_i32, _i32]),
    "gtrap_peakstats": (_int, [_vp] * 5 + [_i32] * 4 + [_vp, _i64, _i32, _vp]),
    "gtrap_cutout": (_int, [_vp] * 6 + [_i32, _i32] + [_vp] * 3 + [_i32, _f64, _i32, _f64, _i32,
                                                                      _vp]),
    "gtrap_blsstats": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _f64, _f64, _vp, _vp, _vp, _vp,
                              _vp, _vp]),
    "gtrap_gtls": (_int, [_vp] * 7 + [_i32] * 5 + [_f32] * 5 + [_int, _vp]),
    "gtrap_tlsgrid": (_int, [_vp, _vp, _i32, _vp, _i32, _vp, _i32, _vp, _i32, _int, _f64, _i32,
                             _vp, _vp, _vp, _vp, _vp]),
    "gtrap_tlsfit_best": (_int, [_vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp]),
    "gtrap_tlsfold_ws_bytes": (_i64, [_i32, _i32]),
    "gtrap_tlsfold": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _i32,
                             _i64, _vp, _i32, _i32, _f64, _i32, _vp, _i64, _i32, _vp]),
    "gtrap_tlsfold_set_tiles": (_int, [_i32, _i32, _i32, _i32]),
    "gtrap_tlsfold_static": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                    _i32, _i64, _vp, _i32, _i32, _f64, _i32, _i32, _vp]),
    "gtrap_tlsfold_prec": (_int, [_i32] + [_vp] * 11 + [_i32, _vp, _vp, _i32, _i64, _vp, _i32,
                                                      _i32, _f64, _i32, _vp, _i64, _i32, _vp]),
    "gtrap_tlsfold_items": (_int, [_vp] * 9 + [_i32] + [_vp] * 6 + [_i32, _i32, _f64, _i32, _vp,
                                                                  _i64, _i32, _vp]),
    "gtrap_tlsstats": (_int, [_vp] * 5 + [_i32] + [_vp] * 4),
}

_lib = None


class GtrapError(RuntimeError):
    """A non-zero status returned by libgtrap_hip.so."""

    def __init__(self, func: str, status: int, text: str):
        super().__init__(f"{func} failed (status {status}): {text}")
        self.status = status


def load() -> C.CDLL:
    """Load the library (once) and declare every signature."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `python -m gtrap_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name: str, *args) -> None:
    """Call an entry point and raise GtrapError on a non-zero status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        txt = lib.gtrap_last_error().decode(errors="replace")
        if st == -1:
            raise ValueError(f"{name}: {txt}")
        raise GtrapError(name, st, txt)


def ptr(buf) -> int | None:
    """Device address of a torch tensor / DeviceAllocation / int, or None."""
    if buf is None:
        return None
    if isinstance(buf, int):
        return buf
    if hasattr(buf, "data_ptr"):
        if not buf.is_cuda:
            raise ValueError("expected a device tensor, got a host tensor")
        return buf.data_ptr()
    if hasattr(buf, "__int__"):
        return int(buf)
    raise TypeError(f"cannot take a device address of {type(buf).__name__}")


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (default: torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


def require_gpu() -> None:
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("gtrap_amd needs a ROCm GPU; none is visible (no CPU fallback)")
    load()
