"""gtrap_amd -- MI355X-native transit search (drop-in for gtrap's GPU calls).

Modules mirroring the reference package (HajimeKawahara/gtrap):

* :mod:`gtrap_amd.gfilter`  -- median-filter detrend (gtrap/gfilter.py)
* :mod:`gtrap_amd.driver`   -- PyCUDA-style mem_alloc / memcpy shim

Kernels live in libgtrap_hip.so (HIP, gfx950), called through the C ABI in
include/gtrap_hip.h.  There is no CPU fallback.
"""
__version__ = "0.1.0"
