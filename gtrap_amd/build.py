"""Build libgtrap_hip.so (gfx950) in-tree with hipcc.

The library is compiled from every ``csrc/*.hip`` and ``csrc/*.cpp`` file into
``gtrap_amd/libgtrap_hip.so``.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgtrap_hip.so"
INCLUDE = PKG.parent / "include"

ARCH = os.environ.get("GTRAP_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is written out as
# fmaf()/fma(), so results are exactly what the source says.
# -packed-fp32-ops is disabled for all device code: some packed FP32 forms
# have given wrong results on this hardware, without any error.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgtrap_hip.so")


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), *HIPFLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")


def build_library(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link ``libgtrap_hip.so``; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    objs = [BUILD / (s.stem + ".o") for s in sources]
    jobs = jobs or min(8, len(sources)) or 1
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in zip(sources, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
