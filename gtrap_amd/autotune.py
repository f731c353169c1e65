"""Autotuner for execution parameters.

Only parameters that do not change results are tuned: the tls_fold kernel's
tile geometry (epochs per staged LDS tile, phase-index cells per sample,
samples per phase bucket, epochs per window-table run), the number of
persistent tls_fold workgroups, the serving step size (work items per step),
and the detrend segment length.  Semantic parameters (bins, widths, L steps,
grids, windows) are never touched: region membership and the best trials do
not depend on the tuned values, and S/N moves by float64 rounding at most
(the tile epochs and the bucket size set the staged ranges and the tiles'
reference epochs; the other parameters leave the bits unchanged).

Results are cached as JSON keyed by device name and workload signature in
``$GTRAP_TUNE_CACHE`` (default ``~/.cache/gtrap_amd/tune.json``).  Lookups
fall back to the table shipped with the package (``tuned.json``, written by
``python -m gtrap_amd.autotune`` on the target GPU), then to the caller's
default.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path

import numpy as np

from . import _native


def cache_path() -> Path:
    p = os.environ.get("GTRAP_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "gtrap_amd" / "tune.json"


def _device_key() -> str:
    """Architecture and CU count (e.g. ``gfx950_256cu``): the marketing name
    depends on whether the runtime finds its id table, so it is not a key."""
    import torch
    prop = torch.cuda.get_device_properties(0)
    arch = getattr(prop, "gcnArchName", "") or torch.cuda.get_device_name(0)
    return f"{arch.split(':')[0].replace(' ', '_')}_{prop.multi_processor_count}cu"


def load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def store(key: str, value) -> None:
    d = load()
    d[key] = value
    p = cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(d, indent=1, sort_keys=True))


PACKAGED = Path(__file__).resolve().with_name("tuned.json")


def lookup(key: str, default=None):
    k = f"{_device_key()}:{key}"
    v = load().get(k)
    if v is None:
        try:
            v = json.loads(PACKAGED.read_text()).get(k)
        except (OSError, ValueError):
            v = None
    return default if v is None else v


def _time(fn, reps=3) -> float:
    import torch
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_tlsfold_blocks(curves, periods_list, widths, nl=10, dtep=0.0204,
                        candidates=(256, 512, 768, 1024, 1536, 2048), save=True) -> dict:
    """Persistent workgroup count for tls_fold on a representative batch."""
    from .tlsfold import Workspace, pack, tlsfold_device
    batch = pack(curves, periods_list)
    ws = Workspace()
    times = {}
    for nb in candidates:
        times[nb] = _time(lambda: tlsfold_device(batch, widths, nl, dtep, nblocks=nb,
                                                 workspace=ws))
    best = min(times, key=times.get)
    res = dict(nblocks=int(best), seconds={str(k): v for k, v in times.items()})
    if save:
        store(f"{_device_key()}:tlsfold.nblocks", int(best))
    return res


TILE_KEY = "tlsfold.tiles"
DEFAULT_TILES = (144, 2, 2, 0)
_tiles_applied = None


def set_tiles(max_tile=144, cells=2, spb=2, te_run=0) -> None:
    """Set the tls_fold tile geometry for the launches that follow."""
    _native.call("gtrap_tlsfold_set_tiles", int(max_tile), int(cells), int(spb), int(te_run))


def apply_tiles(force: bool = False):
    """Apply the tuned tile geometry (cache, then the packaged table, then the
    defaults) once per process; returns it."""
    global _tiles_applied
    if _tiles_applied is None or force:
        v = lookup(TILE_KEY, list(DEFAULT_TILES))
        set_tiles(*v)
        _tiles_applied = tuple(int(a) for a in v)
    return _tiles_applied


def tune_tiles(curves, periods_list, widths, nl=10, dtep=0.0204, nblocks=512,
               max_tiles=(48, 72, 96, 144), cells=(1, 2), spbs=(1, 2, 4), te_runs=(0, 24, 36),
               save=True, log=None) -> dict:
    """Kernel time (HIP events) of tls_fold on a representative batch for
    every tile geometry of the grid; stores the fastest."""
    import torch
    from .tlsfold import Workspace, pack, tlsfold_device
    batch = pack(curves, periods_list)
    ws = Workspace()
    times = {}
    for mt in max_tiles:
        for c in cells:
            for sp in spbs:
                for tr in te_runs:
                    set_tiles(mt, c, sp, tr)

                    def run():
                        a = torch.cuda.Event(enable_timing=True)
                        b = torch.cuda.Event(enable_timing=True)
                        a.record()
                        tlsfold_device(batch, widths, nl, dtep, nblocks=nblocks, workspace=ws)
                        b.record()
                        torch.cuda.synchronize()
                        return a.elapsed_time(b)
                    run()
                    ms = min(run() for _ in range(2))
                    times[f"{mt}/{c}/{sp}/{tr}"] = ms
                    if log:
                        log(f"tiles {mt}/{c}/{sp}/{tr}: {ms:.2f} ms")
    best = min(times, key=times.get)
    geo = [int(v) for v in best.split("/")]
    set_tiles(*geo)
    if save:
        store(f"{_device_key()}:{TILE_KEY}", geo)
    return dict(tiles=geo, ms=times, default_ms=times.get("144/2/2/0"))


def tune_medfilt_seg(n, nq, r, candidates=(256, 512, 1024, 2048, 4096), save=True) -> dict:
    import torch
    from . import gfilter
    x = torch.rand((n, nq), device="cuda") + 0.5
    y = torch.empty_like(x)
    times = {}
    for seg in candidates:
        if seg > n:
            continue
        times[seg] = _time(lambda: gfilter.medfilt_device(x, y, n, nq, r, "div", seg_len=seg))
    best = min(times, key=times.get)
    if save:
        store(f"{_device_key()}:medfilt.seg.{r}.{nq >= 64}", int(best))
    return dict(seg_len=int(best), seconds={str(k): v for k, v in times.items()})


def tune_serve(trace_fn, candidates=((16384, 1024), (32768, 1024), (65536, 1024),
                                     (32768, 2048), (65536, 2048)), save=True) -> dict:
    """Serving step size and workgroups: served light curves/s on a trace
    (``trace_fn()`` returns fresh requests, all available at once)."""
    from .serve import ServerConfig, TransitServer
    import torch
    rates = {}
    for ips, nb in candidates:
        srv = TransitServer(ServerConfig(items_per_step=ips, nblocks=nb))
        srv.run(trace_fn(), realtime=False)  # warm (graph capture)
        reqs = trace_fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        srv.run(reqs, realtime=False)
        torch.cuda.synchronize()
        rates[f"{ips}x{nb}"] = len(reqs) / (time.perf_counter() - t0)
        del srv
    best = max(rates, key=rates.get)
    ips, nb = (int(v) for v in best.split("x"))
    if save:
        store(f"{_device_key()}:serve.items_per_step", ips)
        store(f"{_device_key()}:serve.nblocks", nb)
    return dict(items_per_step=ips, nblocks=nb, rates=rates)


def main(argv=None) -> int:
    """Tune the fold kernel's tile geometry and workgroups and the serving
    step on the current GPU; prints the choices and stores them in the
    cache."""
    import argparse
    from . import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=48)
    args = ap.parse_args(argv)
    _native.require_gpu()
    reqs, _ = synth.serving_trace(24, seed=7)
    out = dict(device=_device_key())
    big = sorted(reqs, key=lambda r: len(r.t))[-12:]
    out["tiles"] = tune_tiles([(r.t, r.flux) for r in big], [r.periods for r in big],
                              np.linspace(0.04, 0.3, 10),
                              log=lambda m: print(m, flush=True))
    print(json.dumps({k: v for k, v in out["tiles"].items() if k != "ms"}), flush=True)
    out["tlsfold"] = tune_tlsfold_blocks([(r.t, r.flux) for r in reqs[:6]],
                                         [r.periods for r in reqs[:6]],
                                         np.linspace(0.04, 0.3, 10),
                                         candidates=(256, 512, 1024, 2048))
    print(json.dumps(out["tlsfold"]), flush=True)
    out["serve"] = tune_serve(lambda: synth.serving_trace(args.requests, seed=8)[0],
                              candidates=((16384, 512), (32768, 512), (65536, 512),
                                          (16384, 1024), (32768, 1024), (65536, 1024),
                                          (131072, 1024)))
    print(json.dumps(out["serve"]), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
