"""Benchmark: one JSON line on stdout (Measurement).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Only one GPU is used; --gpus N with N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    cores = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


# ----------------------------------------------------------------------------
# gfilter workload: 4096 curves x 65536 samples, windows r = 7, 24, 64, divide
# ----------------------------------------------------------------------------
def bench_gfilter(steps: int, warmup: int):
    import torch
    from gtrap_amd import gfilter, synth
    n, nq = 65536, 4096
    radii = (7, 24, 64)
    log(f"gfilter: generating {nq}x{n} image on device")
    img = synth.config3(nq=nq, n=n, device="cuda")
    out = torch.empty_like(img)
    torch.cuda.synchronize()
    ev = {r: [] for r in radii}

    def step(record):
        for r in radii:
            if record:
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
            gfilter.medfilt_device(img, out, n, nq, r, "div")
            if record:
                b.record()
                ev[r].append((a, b))

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = {r: float(np.mean([a.elapsed_time(b) for a, b in ev[r]])) for r in radii}
    log("gfilter kernel ms: " + ", ".join(f"r={r}: {kms[r]:.3f}" for r in radii))
    bytes_per = 8.0 * n * nq
    dom = max(radii, key=lambda r: kms[r])
    ach = bytes_per / (kms[dom] * 1e-3) / 1e9
    # CPU baseline: scipy.signal.medfilt on a bounded sample of curves
    from scipy.signal import medfilt
    sample = img[:, :8].cpu().numpy()
    cores, model = cpu_info()
    t1 = time.perf_counter()
    nc = 0
    while time.perf_counter() - t1 < 10.0 and nc < sample.shape[1]:
        x = sample[:, nc]
        for r in radii:
            x / medfilt(x, 2 * r + 1)
        nc += 1
    cpu_rate = nc / (time.perf_counter() - t1)
    res = dict(
        metric="light curves/s (gfilter median detrend, 3 windows)",
        value=nq * steps / wall, unit="light curves/s", ms_per_step=wall / steps * 1e3,
        dtype="f32", config=dict(workload="gfilter 4096 curves x 65536 samples, r=7,24,64, div",
                                 n=n, nq=nq, radii=list(radii)),
        roofline=dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=ach / HBM_PEAK_GBS, traffic=None,
                      kernel=f"medfilt_win_kernel r={dom}",
                      kernel_ms={str(r): kms[r] for r in radii}),
        cpu_baseline=dict(value=cpu_rate, unit="light curves/s", cores=1, kind="port",
                          sample=f"{nc} curves x 65536 samples, scipy.signal.medfilt r=7,24,64 "
                                 f"then divide, 1 thread ({model})"),
    )
    return res


WORKLOADS = {"gfilter": bench_gfilter}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="gfilter", choices=sorted(WORKLOADS))
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the transit search shards by curve with no exchange; see DESIGN.md)",
              file=sys.stderr)
        return 2
    import torch
    from gtrap_amd import _native
    _native.require_gpu()
    torch.cuda.init()
    res = WORKLOADS[args.workload](args.steps, args.warmup)
    line = dict(metric=res.pop("metric"), value=res.pop("value"), unit=res.pop("unit"),
                n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=res.pop("ms_per_step"), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype=res.pop("dtype"), data="synthetic", **res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
