"""gtrap_amd -- MI355X-native transit search (drop-in for gtrap's GPU calls).

Modules mirroring the reference package (HajimeKawahara/gtrap), same names,
argument order and meaning:

* :mod:`gtrap_amd.gfilter`        -- median-filter detrend, offset (gtrap/gfilter.py)
* :mod:`gtrap_amd.geebls`         -- box least squares (gtrap/geebls.py, include/geebls.h)
* :mod:`gtrap_amd.geebls_simple`  -- BLS without the offset (gtrap/geebls_simple.py)
* :mod:`gtrap_amd.getstat`        -- BLS periodogram statistics (gtrap/getstat.py)
* :mod:`gtrap_amd.gtls`           -- single-transit trapezoid search, fast (float64
  prefix sums) and parity (float32 scans) arithmetic (gtrap/gtls.py,
  include/gtls_simple.h, include/gtls_absolute.h)
* :mod:`gtrap_amd.tls`            -- tls.py's trapezoid fit on a device grid (gtrap/tls.py)
* :mod:`gtrap_amd.driver`         -- the pycuda.driver calls the reference's drivers use
  (mem_alloc, memcpy_htod, memcpy_dtoh)

Beyond the reference's modules:

* :mod:`gtrap_amd.tlsfold`    -- periodic trapezoid search over a period grid
  (tls.py's statistic on folded time; an opt-in least-squares detection statistic)
* :mod:`gtrap_amd.serve`      -- continuous-batching server for ragged TLS+detrend requests
* :mod:`gtrap_amd.peakstats`  -- peak statistics of the S/N series (gtrap/detect_peaks.py,
  gtrap/gtls.py:244-275)
* :mod:`gtrap_amd.image`      -- light-curve image builder (gtrap/read_keplerlc.py)
* :mod:`gtrap_amd.picktrap`   -- cut-outs around a detection (gtrap/picktrap.py)
* :mod:`gtrap_amd.inject`     -- synthetic noise and trapezoid injection (gtrap/ggenmock.py)
* :mod:`gtrap_amd.precision`  -- arithmetic variants of BLS and tls_fold (precision study)
* :mod:`gtrap_amd.autotune`   -- launch geometry tuner (``tuned.json``)
* :mod:`gtrap_amd.synth`      -- synthetic workloads of the benchmarks and tests

Kernels live in libgtrap_hip.so (HIP, gfx950), called through the C ABI in
include/gtrap_hip.h.  There is no CPU fallback: every compute call raises when
the library or the GPU is missing.
"""
__version__ = "0.2.0"
