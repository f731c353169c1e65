"""Median-detrend oracles (gtrap/gfilter.py:13-85).

* :func:`detrend_scipy` -- ``x - medfilt(x, 2r+1)`` (or ``x / medfilt``) per
  curve with ``scipy.signal.medfilt``, the zero-padded median the reference
  kernel computes (FILVAL 0.0 halo, gfilter.py:12,25-36).  Library oracle.
* :func:`gfilter_bisect` -- C restatement of the reference's bisection
  selector itself (gfilter.py:40-76), used to pin that the reference's
  selector and scipy's median agree.
* :func:`detrend_sort` -- brute force with ``np.sort`` (tiny inputs).

Pinned by tests/test_oracle_medfilt.py: scipy == bisection == brute force
on random, tied, constant and edge-dominated inputs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._clib import fptr, lib


def _apply(x: np.ndarray, med: np.ndarray, mode: str) -> np.ndarray:
    x = x.astype(np.float32)
    med = med.astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        return (x - med) if mode == "sub" else (x / med)


def median_scipy(img: np.ndarray, r: int) -> np.ndarray:
    """Zero-padded window median along axis 0 of an (n, nq) float32 image."""
    from scipy.signal import medfilt
    img = np.asarray(img, dtype=np.float32)
    out = np.empty_like(img)
    for q in range(img.shape[1]):
        out[:, q] = medfilt(img[:, q], 2 * r + 1)
    return out


def detrend_scipy(img: np.ndarray, r: int, mode: str = "sub") -> np.ndarray:
    img = np.asarray(img, dtype=np.float32)
    return _apply(img, median_scipy(img, r), mode)


def detrend_sort(img: np.ndarray, r: int, mode: str = "sub") -> np.ndarray:
    img = np.asarray(img, dtype=np.float32)
    n, nq = img.shape
    pad = np.zeros((n + 2 * r, nq), dtype=np.float32)
    pad[r:r + n] = img
    med = np.empty_like(img)
    for i in range(n):
        med[i] = np.sort(pad[i:i + 2 * r + 1], axis=0)[r]
    return _apply(img, med, mode)


def gfilter_bisect(img: np.ndarray, r: int, mode: str = "sub") -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float32)
    n, nq = img.shape
    out = np.empty_like(img)
    f = lib().orc_gfilter_bisect
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    f(fptr(img), fptr(out), n, nq, r, 0 if mode == "sub" else 1)
    return out
