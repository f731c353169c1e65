"""Benchmark: one JSON line on stdout.

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
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12  # 128 B/clk per CU, 256 CUs, 2.4 GHz
FP64_PEAK = 78.6  # TFLOP/s, MI355X dense FP64
# FP64 operations of one (epoch, W, L) trial from the tile prefix sums: the
# region sums from prefix differences, shifted to epoch coordinates for the
# ingress and egress regions (29), and the closed-form trapezoid fit,
# residual and rank key sign(S/N)*(S/N)^2 (74); +, -, *, / count one each,
# a fused multiply-add two
TRIAL_FLOPS = 103
# HBM bytes of the serving steps' fold per evaluated trial, measured offline:
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch of
# tlsfold_kernel<true> over the default serve workload, scaled by the same
# counters over a 1 GiB copy (factors 2.0 and 1.0): 13.69 B read + 14.73 B
# written per evaluated trial.  The line's traffic is this times the run's
# evaluated trials per launch (launches split the trace differently from run
# to run, so a per-launch constant would not carry over).
SERVE_FOLD_BYTES_PER_TRIAL = 28.42
SERVE_FOLD_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this workload "
                             "of the same library (separate runs, not measured in this one): "
                             "28.42 HBM bytes per evaluated trial times this run's evaluated "
                             "trials per launch, profiles/tlsfold_serve_hbm_traffic.md")


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
    img = synth.detrend_image(nq=nq, n=n, device="cuda")
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
    # the wide windows of the reference's drivers (r = 100, 128, 512) on the
    # first 1024 curves
    wide = {}
    for r in (100, 128, 512):
        gfilter.medfilt_device(img, out, n, nq, r, "div", cols_out=1024)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        gfilter.medfilt_device(img, out, n, nq, r, "div", cols_out=1024)
        b.record()
        torch.cuda.synchronize()
        wide[str(r)] = dict(kernel_ms=a.elapsed_time(b), curves=1024,
                            light_curves_per_s=1024 / (a.elapsed_time(b) * 1e-3))
    log("gfilter wide windows (1024 curves): " +
        ", ".join(f"r={k}: {v['kernel_ms']:.2f} ms" for k, v in wide.items()))
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
                      kernel=(f"medfilt_block_kernel r={dom}" if dom in (24, 64)
                              else f"medfilt_win_kernel r={dom}"),
                      kernel_ms={str(r): kms[r] for r in radii}),
        wide_windows=wide,
        cpu_baseline=dict(value=cpu_rate, unit="light curves/s", cores=1, kind="port",
                          sample=f"{nc} curves x 65536 samples, scipy.signal.medfilt r=7,24,64 "
                                 f"then divide, 1 thread ({model})"),
    )
    return res


# ----------------------------------------------------------------------------
# geebls workload: 1024 curves x 16384 samples, 2048 frequencies, nb 1024
# ----------------------------------------------------------------------------
def bench_geebls(steps: int, warmup: int):
    import torch
    from gtrap_amd import geebls, getstat, synth
    log("geebls: generating 1024 curves x 16384 samples")
    d = synth.bls_batch()
    g = synth.bls_batch_grid()
    lc, tu, ntrue = d["lc"], d["tu"], d["ntrue"]
    n, nq = lc.shape
    nf = g["nf"]
    off = (np.where(tu >= 0, lc, 0).astype(np.float64).sum(0) / n).astype(np.float32)
    dl, dt, do, dn = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (lc, tu, off, ntrue))
    blsp = torch.empty(nq * nf, dtype=torch.float32, device="cuda")
    p1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    evs = []

    def step(record):
        if record:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
        geebls.geebls_device(blsp, p1, p2, dl, dt, do, dn, n, nq, nf, g["nb"], g["kma"], g["kmi"],
                             g["kkmi"], g["fmin"], g["df"], nthread=g["nthread"])
        if record:
            b.record()
            evs.append((a, b))
        getstat.blsstats_device(blsp, p1, p2, nf, nq, g["df"], g["fmin"])

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    trials = nf * nq * g["nb"] * (g["kma"] - g["kmi"])
    lds_bytes = nf * nq * (8.0 * n + 8.0 * (g["nb"] + g["kma"]))
    log(f"geebls kernel {kms:.3f} ms; trial evals/s {trials / kms * 1e3:.3e}")
    cores, model = cpu_info()
    from oracle import bls as ob
    t1 = time.perf_counter()
    nfs = 0
    while time.perf_counter() - t1 < 10.0 and nfs < nf:
        ob.geebls(lc, tu, off, ntrue, nf, g["nb"], g["kma"], g["kmi"], g["kkmi"], g["fmin"],
                  g["df"], nthread=g["nthread"], qs=(0, 1), fs=(nfs, nfs + 16))
        nfs += 16
    cpu_s_per_curve = (time.perf_counter() - t1) / nfs * nf
    ach = lds_bytes / (kms * 1e-3) / 1e12
    lds_peak = LDS_PEAK_TBS
    return dict(
        metric="light curves/s and trial evals/s (geebls + getstat)",
        value=nq * steps / wall, unit="light curves/s", ms_per_step=wall / steps * 1e3,
        dtype="f32", trial_evals_per_s=trials / (kms * 1e-3),
        config=dict(workload="batched geebls 1024 curves x 16384 samples, 2048 frequencies, "
                             "nb 1024, 12 widths", n=n, nq=nq, nf=nf, nb=g["nb"]),
        roofline=dict(bound="lds", achieved=ach, peak=lds_peak, unit="TB/s",
                      frac=ach / lds_peak, traffic=None, kernel="bls_kernel", kernel_ms=kms),
        cpu_baseline=dict(value=1.0 / cpu_s_per_curve, unit="light curves/s", cores=1,
                          kind="port", sample=f"{nfs} frequencies of 1 curve with the C oracle, "
                                              f"scaled to {nf} ({model})"),
    )


def cpu_trial_rates(t, x, periods, widths, nl, dtep, cap=8.0):
    """CPU baselines of the trapezoid search (the C oracle and the numpy
    restatement of tls.py, one core and a process pool)."""
    from oracle import timing
    return timing.trial_rates(t, x, periods, widths, nl, dtep, cap=cap, log=log)


# ----------------------------------------------------------------------------
# serving workload: TLS+detrend requests, served vs plain
# ----------------------------------------------------------------------------
SERVE_REQUESTS = int(os.environ.get("GTRAP_BENCH_REQUESTS", "160"))


def bench_serve(steps: int, warmup: int):
    import torch
    from gtrap_amd import synth
    from gtrap_amd.serve import ServerConfig, TransitServer, search_plain, PlainSearch
    cfg = ServerConfig.tuned()
    log(f"serve: building traces of {SERVE_REQUESTS} requests; items/step {cfg.items_per_step}, "
        f"workgroups {cfg.nblocks}")
    traces = [synth.serving_trace(SERVE_REQUESTS, seed=1004 + k)
              for k in range(warmup + steps)]
    srv = TransitServer(cfg)
    srv.fold_log = []
    # warm-up (graph capture, allocator, code objects)
    for k in range(warmup):
        srv.run(traces[k][0], realtime=True)
    torch.cuda.synchronize()
    # served, Poisson arrivals in real time
    walls, lat, trials, nreq = [], [], 0, 0
    rec, ninj = 0, 0
    last = None
    for k in range(warmup, warmup + steps):
        reqs, truth = traces[k]
        trials += sum(synth.trial_count(r) for r in reqs)
        nreq += len(reqs)
        t0 = time.perf_counter()
        res = srv.run(reqs, t0=t0, realtime=True)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        lat += [r.latency for r in res]
        for r in res:
            if truth[r.rid] is not None:
                ninj += 1
                rec += synth.recovered(r.best_period, truth[r.rid])
        last = {r.rid: r for r in res}
        log(f"served trace {k}: {len(res)} requests in {walls[-1]:.3f} s, {srv.steps} steps")
    wall = sum(walls)
    served = nreq / wall
    lat = np.array(lat)
    # plain: one launch sequence per curve over the same requests as the last
    # served trace (every request of it)
    ctx = PlainSearch()
    sub = traces[-1][0]
    search_plain(sub[0], cfg.widths, cfg.nl, cfg.dtep, cfg.r, ctx=ctx)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plat = []
    same = True
    for r in sub:
        a = time.perf_counter()
        p = search_plain(r, cfg.widths, cfg.nl, cfg.dtep, cfg.r, ctx=ctx)
        plat.append(time.perf_counter() - a)
        g = last[r.rid]
        same &= (p.best_index, p.w, p.l, p.e, p.sn) == (g.best_index, g.w, g.l, g.e, g.sn)
    pwall = time.perf_counter() - t0
    plain = len(sub) / pwall
    ptrials = sum(synth.trial_count(r) for r in sub)
    served_last = len(sub) / walls[-1]
    log(f"plain: {len(sub)} requests in {pwall:.3f} s (served the same in {walls[-1]:.3f} s); "
        f"results identical: {same}")
    # detection with the least-squares statistic on the last traces (served
    # again, untimed)
    import dataclasses
    srv_d = TransitServer(dataclasses.replace(cfg, stat="dchi2"))
    rec_d = ninj_d = 0
    ndet = min(steps, 8)
    for k in range(warmup + steps - ndet, warmup + steps):
        reqs, truth = traces[k]
        for r in srv_d.run(reqs, realtime=False):
            if truth[r.rid] is not None:
                ninj_d += 1
                rec_d += synth.recovered(r.best_period, truth[r.rid])
    log(f"detection: tls.py S/N {rec}/{ninj}, least-squares statistic {rec_d}/{ninj_d}")
    # dominant kernel: the serving steps' tlsfold (item-list instantiation),
    # every launch of the run (warm-up included, as a profiler sees them)
    # bracketed by HIP events on the serving stream
    nlaunch, kms, ktr, ksec = srv.fold_stats()
    # trials with L = 0 are NaN by construction and are not evaluated
    evaluated = ktr * (cfg.nl - 1) / cfg.nl
    ach = TRIAL_FLOPS * evaluated / ksec / 1e12
    # CPU baselines on one mid-size request of the last trace (scipy detrend +
    # the trapezoid search), scaled by trial evaluations to the mean request
    from oracle import medfilt as om
    cores, model = cpu_info()
    r0 = sorted(sub, key=lambda r: len(r.t))[len(sub) // 2]
    t1 = time.perf_counter()
    xd = om.detrend_scipy(r0.flux[:, None], cfg.r, "div")[:, 0]
    det_s = time.perf_counter() - t1
    rates = cpu_trial_rates(r0.t, xd, r0.periods, cfg.widths, cfg.nl, cfg.dtep)
    # light curves/s of the request's size class, scaled by trial evaluations
    # to the trace's mean request
    scale = synth.trial_count(r0) / (trials / nreq)
    lcs = lambda curves: 1.0 / (scale / curves + det_s)
    return dict(
        metric="light curves/s and trial evals/s (TLS+detrend), served vs plain; p99 latency",
        value=served, unit="light curves/s", ms_per_step=wall / steps * 1e3, dtype="f64",
        trial_evals_per_s=trials / wall,
        plain=dict(value=plain, unit="light curves/s", trial_evals_per_s=ptrials / pwall,
                   requests=len(sub), same_requests_as="the last served trace",
                   results_identical=bool(same), p50_ms=float(np.median(plat) * 1e3),
                   p99_ms=float(np.percentile(plat, 99) * 1e3)),
        served_vs_plain=dict(light_curves=served_last / plain,
                             trial_evals=(ptrials / walls[-1]) / (ptrials / pwall),
                             target=10.0, met=bool(served_last / plain >= 10.0)),
        latency_ms=dict(p50=float(np.percentile(lat, 50) * 1e3),
                        p99=float(np.percentile(lat, 99) * 1e3),
                        max=float(lat.max() * 1e3),
                        note="arrival to result of requests arriving in 160-request bursts "
                             "(80 ms at 2000 req/s) that drain before the next burst; "
                             "open-loop figures: --workload openloop"),
        detection=dict(injected=ninj, recovered=rec, fraction=rec / max(ninj, 1),
                       rule="best period within 1% of P, P/2, P/3, 2P or 3P",
                       statistic="tls.py S/N (the served default)",
                       least_squares=dict(injected=ninj_d, recovered=rec_d,
                                          fraction=rec_d / max(ninj_d, 1),
                                          traces=ndet, statistic="stat=dchi2")),
        config=dict(workload="serving ragged TLS+detrend: Poisson 2000 req/s, n log-uniform "
                             "2048-70000 at 0.0204 d, periods geomspace(0.5, baseline/3, "
                             "500-4000), 10 W x 10 L, detrend r=64 divide",
                    requests_per_step=SERVE_REQUESTS, items_per_step=cfg.items_per_step,
                    nblocks=cfg.nblocks, hbm_budget_gib=cfg.hbm_budget / 2 ** 30),
        roofline=dict(bound="fp64-valu", achieved=ach, peak=FP64_PEAK, unit="TFLOP/s",
                      frac=ach / FP64_PEAK,
                      traffic=SERVE_FOLD_BYTES_PER_TRIAL * evaluated / nlaunch,
                      traffic_source=SERVE_FOLD_TRAFFIC_SOURCE,
                      kernel="tlsfold_kernel<true> (serving step)", launches=nlaunch,
                      kernel_ms=kms, trials_per_launch=ktr / max(nlaunch, 1),
                      evaluated_trials_per_launch=evaluated / max(nlaunch, 1),
                      flops_per_trial=TRIAL_FLOPS),
        cpu_baseline=dict(value=lcs(rates["c_allcores_curves"]), unit="light curves/s",
                          cores=rates["cores"], kind="port",
                          trial_evals_per_s=rates["c_allcores"],
                          one_core=dict(light_curves_per_s=lcs(rates["c_1core_curves"]),
                                        trial_evals_per_s=rates["c_1core"]),
                          numpy_tls_py=dict(one_core_trial_evals_per_s=rates["numpy_1core"],
                                            all_cores_trial_evals_per_s=rates["numpy_allcores"]),
                          sample=f"random periods of one n={len(r0.t)} request "
                                 f"({len(r0.periods)} periods), about 8 s per leg: the C oracle "
                                 f"(direct window sums) on {rates['cores']} processes and on 1 "
                                 f"core, and the numpy restatement of tls.py's statistic "
                                 f"(vectorised over epochs); curves/s = periods/s over the grid, "
                                 f"plus scipy medfilt, scaled by trial evaluations to the "
                                 f"trace's mean request ({model})"),
    )


# ----------------------------------------------------------------------------
# open-loop serving: a Poisson stream for a fixed duration at several rates
# ----------------------------------------------------------------------------
OPENLOOP_RATES = tuple(float(v) for v in os.environ.get(
    "GTRAP_OPENLOOP_RATES", "60,120,180,2000").split(","))
OPENLOOP_SECONDS = float(os.environ.get("GTRAP_OPENLOOP_SECONDS", "10"))


def bench_openloop(steps: int, warmup: int):
    """Arrivals at each rate for OPENLOOP_SECONDS (default 10 s), served until
    drained: queueing latency p50/p99 from arrival to result, the completion
    rate, and the highest rate whose p99 stays below 2 s."""
    import torch
    from gtrap_amd import synth
    from gtrap_amd.serve import ServerConfig, TransitServer
    cfg = ServerConfig.tuned()
    srv = TransitServer(cfg)
    warm, _ = synth.serving_trace(40, seed=3000)
    srv.run(warm, realtime=False)
    torch.cuda.synchronize()
    rows = []
    for lam in OPENLOOP_RATES:
        reqs, truth = synth.serving_stream(lam, OPENLOOP_SECONDS, seed=2004 + int(lam))
        log(f"openloop: {len(reqs)} requests at {lam:g} req/s over {OPENLOOP_SECONDS:g} s")
        t0 = time.perf_counter()
        res = srv.run(reqs, t0=t0, realtime=True,
                      progress=lambda d, q: log(f"  {d} done, {q} queued"))
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in res])
        rows.append(dict(rate=lam, requests=len(res), wall_s=wall,
                         completion_rate=len(res) / wall,
                         p50_ms=float(np.percentile(lat, 50) * 1e3),
                         p99_ms=float(np.percentile(lat, 99) * 1e3),
                         max_ms=float(lat.max() * 1e3)))
        log(f"  {rows[-1]}")
    ok = [r["rate"] for r in rows if r["p99_ms"] < 2000.0]
    top = rows[-1]
    return dict(metric="open-loop serving: queueing latency p50/p99 and completion rate",
                value=top["completion_rate"], unit="light curves/s",
                ms_per_step=top["wall_s"] * 1e3, dtype="f64", rates=rows,
                sustained_rate=max(ok) if ok else None,
                config=dict(workload=f"open-loop Poisson streams of TLS+detrend requests for "
                                     f"{OPENLOOP_SECONDS:g} s at "
                                     f"{', '.join(f'{r:g}' for r in OPENLOOP_RATES)} req/s, "
                                     f"served until drained", seconds=OPENLOOP_SECONDS),
                roofline=None, cpu_baseline=None)


# ----------------------------------------------------------------------------
# single-curve workload: 4096 samples, 1000 periods 0.5-20 d, 10 W x 10 L
# ----------------------------------------------------------------------------
def bench_tls1(steps: int, warmup: int):
    import torch
    from gtrap_amd import synth, tlsfold
    d = synth.tls_single_curve()
    t, x, per, widths, nl = d["t"], d["flux"], d["periods"], d["widths"], d["nl"]
    batch = tlsfold.pack([(t, x)], [per])
    for _ in range(max(1, warmup)):
        tlsfold.tlsfold_device(batch, widths, nl, synth.CADENCE)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    reps = max(1, steps) * 20
    t0 = time.perf_counter()
    a.record()
    for _ in range(reps):
        out = tlsfold.tlsfold_device(batch, widths, nl, synth.CADENCE)
    b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = a.elapsed_time(b) / reps
    sn = out["sn"].cpu().numpy()
    from oracle import tlsfold as ot
    best, bval, sde = ot.tlsstats(sn)
    snd = tlsfold.tlsfold_device(batch, widths, nl, synth.CADENCE, stat="dchi2")["sn"]
    bd, vd, sded = ot.tlsstats(snd.cpu().numpy())
    trials = int(np.sum(np.ceil(per / synth.CADENCE))) * len(widths) * nl
    evaluated = trials * (nl - 1) / nl
    ach = TRIAL_FLOPS * evaluated / (kms * 1e-3) / 1e12
    log(f"tls1: {kms:.3f} ms per curve; best period {per[best]:.4f} d (injected "
        f"{d['truth']['P']} d), S/N {bval:.3f}, SDE {sde:.2f}")
    rates = cpu_trial_rates(t, x, per, widths, nl, synth.CADENCE, cap=6.0)
    return dict(metric="light curves/s and trial evals/s (plain TLS, one curve)",
                value=1e3 / kms, unit="light curves/s", ms_per_step=wall / reps * 1e3,
                dtype="f64", trial_evals_per_s=trials / (kms * 1e-3),
                detection=dict(best_period=float(per[best]), injected_period=d["truth"]["P"],
                               recovered=bool(synth.recovered(per[best], d["truth"])),
                               sn=float(bval), sde=float(sde), statistic="tls.py S/N",
                               least_squares=dict(best_period=float(per[bd]),
                                                  recovered=bool(synth.recovered(per[bd],
                                                                                 d["truth"])),
                                                  value=float(vd), sde=float(sded))),
                config=dict(workload="plain TLS: 1 curve x 4096 samples at 0.0204 d, trapezoid "
                                     "depth 2e-3 at P 3.7 d, 1000 periods 0.5-20 d, 10 W x 10 L",
                            n=len(t), periods=len(per)),
                roofline=dict(bound="fp64-valu", achieved=ach, peak=FP64_PEAK, unit="TFLOP/s",
                              frac=ach / FP64_PEAK, traffic=None, kernel="tlsfold_kernel<false>",
                              kernel_ms=kms, flops_per_trial=TRIAL_FLOPS),
                cpu_baseline=dict(value=rates["c_allcores_curves"], unit="light curves/s",
                                  cores=rates["cores"], kind="port",
                                  trial_evals_per_s=rates["c_allcores"],
                                  one_core=rates["c_1core_curves"],
                                  numpy_tls_py=dict(one_core=rates["numpy_1core_curves"],
                                                    all_cores=rates["numpy_allcores_curves"]),
                                  sample="random periods of the grid, about 6 s per leg: the C "
                                         "oracle on a process pool and on 1 core, and the "
                                         "numpy restatement of tls.py likewise (light "
                                         "curves/s = periods/s over the 1000-period grid)"))


# ----------------------------------------------------------------------------
# precision workload: BLS arithmetic variants against float64
# ----------------------------------------------------------------------------
def bench_precision(steps: int, warmup: int):
    """256 curves x 262144 samples (2-minute cadence, 364 d), depth 1e-4 at
    SNR 7-15, 2048 resolved frequencies around the transit period: SR and SDE
    errors of each arithmetic variant against the float64 one."""
    from gtrap_amd import precision, synth
    nf = int(os.environ.get("GTRAP_BENCH_PREC_NF", "2048"))
    log(f"precision: generating 256 curves x 262144 samples, {nf} frequencies")
    d = synth.precision_batch(nf=nf)
    t0 = time.perf_counter()
    res = precision.study(d)
    wall = time.perf_counter() - t0
    for k, v in res.items():
        log(f"{k}: {v}")
    # the periodic trapezoid search under the same axes: 32 full-length curves,
    # 32 periods around the transit period (resolving every transit)
    t32 = d["t64"].astype(np.float32)
    ntc = 32
    curves = [(t32, d["lc"][:, q]) for q in range(ntc)]
    per = d["truth"]["P"] + (np.arange(32) - 16) * 5e-4
    tls_res = precision.tls_study(curves, [per] * ntc, np.linspace(0.04, 0.3, 10), 10,
                                  2.0 / 1440.0)
    for k, v in tls_res.items():
        log(f"tls {k}: {v}")
    own = res["f32/exact32/f32"]
    g = d["grid"]
    n, nq = d["lc"].shape
    folds = float(nf) * nq * n
    kms = own["ms"]
    lds_bytes = nf * nq * (12.0 * n + 8.0 * (g["nb"] + g["kma"]))
    ach = lds_bytes / (kms * 1e-3) / 1e12
    lds_peak = LDS_PEAK_TBS
    return dict(
        metric="precision study: SR / SDE error vs float64 (geebls arithmetic variants)",
        value=own["sde_rel_max"], unit="max relative SDE error (gtrap_geebls arithmetic)",
        ms_per_step=wall * 1e3, higher_is_better=False, dtype="f32",
        variants=res,
        tls_variants=dict(results=tls_res, workload="tls_fold on 32 of the curves (262144 "
                                                     "samples, 364 d), 32 periods around P, "
                                                     "10 W x 10 L, epochs every 2 min"),
        config=dict(workload="precision: 256 curves x 262144 samples at 2 min, depth 1e-4, "
                             f"SNR 7-15, {nf} frequencies, nb 1024, kma 23",
                    n=n, nq=nq, nf=nf, nb=g["nb"]),
        roofline=dict(bound="lds", achieved=ach, peak=lds_peak, unit="TB/s", frac=ach / lds_peak,
                      traffic=None, kernel="blsprec_kernel f32/exact32/f32", kernel_ms=kms,
                      sample_folds_per_s=folds / (kms * 1e-3)),
        cpu_baseline=None)


# ----------------------------------------------------------------------------
# copy workload: a plain device copy of known size (HBM counter calibration)
# ----------------------------------------------------------------------------
COPY_BYTES = 1 << 30


def bench_copy(steps: int, warmup: int):
    """Copies a 1 GiB float32 buffer ``steps`` times: a profiler's
    FETCH_SIZE / WRITE_SIZE per dispatch should read 1 GiB each."""
    import torch
    n = COPY_BYTES // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    for _ in range(steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = a.elapsed_time(b) / steps
    gbs = 2.0 * COPY_BYTES / (kms * 1e-3) / 1e9
    return dict(metric="device copy bandwidth (read + write)", value=gbs, unit="GB/s",
                ms_per_step=wall / steps * 1e3, dtype="f32",
                config=dict(workload="copy of one 1 GiB float32 buffer", bytes=COPY_BYTES),
                roofline=dict(bound="hbm", achieved=gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=gbs / HBM_PEAK_GBS, traffic=2.0 * COPY_BYTES,
                              kernel="torch copy", kernel_ms=kms),
                cpu_baseline=None)


# ----------------------------------------------------------------------------
# gtls workload: the single-transit-epoch search at the reference driver's
# Kepler setting (kepler/gtls_mockkepler.py:240-260): 72000 samples at
# 0.0204 d, nw = 1024 widths over 1-6 d, nl = 20, nsc = int(wmax/deltat + 3)
# ----------------------------------------------------------------------------
GTLS_NQ = 4


# FP64 operations of one (epoch, W, L) trial of the fast gtls kernel from the
# tile prefix sums: region sums (prefix differences, the B/D moments moved to
# epoch coordinates) 33, region-A index quirk 5, centred sums 17, denA 16,
# numB 11, H, W/4 and L^2 4, the expanded residual times L^2 36, the rank
# comparison with the best trial 6; +, -, *, / one each.  For the epoch that
# is its tile's reference the region B and D moments need no shift to epoch
# coordinates (18 operations fewer); with one epoch per tile (nw > 512) that
# is every trial.
GTLS_TRIAL_FLOPS = 128
GTLS_SHIFT_FLOPS = 18


def gtls_trial_flops(nw: int) -> float:
    """FP64 operations per trial of the fast gtls kernel, averaged over the
    epochs of a tile (one of its E epochs is the reference)."""
    e = max(1, min(64, 1024 // nw))
    return GTLS_TRIAL_FLOPS - GTLS_SHIFT_FLOPS / e


def bench_gtls(steps: int, warmup: int):
    import torch
    from gtrap_amd import gtls
    n, nq, nw, nl = 72000, GTLS_NQ, 1024, 20
    wmin, wmax = 1.0, 6.0
    rng = np.random.default_rng(1006)
    t = np.arange(n) * 0.0204
    tu = np.repeat(t[:, None], nq, 1)
    tu[rng.random((n, nq)) < 0.02] = -1000.0
    x = 1.0 + 1e-3 * rng.standard_normal((n, nq))
    from gtrap_amd import synth
    for q in range(nq):  # one transit-like dip per curve
        x[:, q] -= synth.trapezoid_dip(t, rng.uniform(100, 1300), None, 3e-3, 0.4, 0.04)
    lc = (2.0 - x).astype(np.float32)  # transit flip, kepler/gtls_mockkepler.py:213-215
    tu = tu.astype(np.float32)
    gp = gtls.gtls_params(tu[:, 0], wmin, wmax, nw, nl)
    deltat, dw, nsc = float(gp["deltat"]), float(gp["dw"]), int(gp["nsc"])
    dl, dt = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (lc, tu))
    outs = [torch.empty(n * nq, dtype=torch.float32, device="cuda") for _ in range(5)]
    log(f"gtls: {nq} curves x {n} samples, nw {nw}, nl {nl}, nsc {nsc}")

    def timed(arith, reps):
        def run():
            gtls.gtls_device(*outs, dl, dt, n, nq, nw, nl, nsc, wmax, dw, deltat, arith=arith)
        for _ in range(max(1, warmup)):
            run()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record()
        for _ in range(reps):
            run()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps, (time.perf_counter() - t0) / reps

    kms, wall = timed("fast", steps)
    fast_out = [o.cpu().numpy() for o in outs]
    pms, _ = timed("parity", 1)
    par_out = [o.cpu().numpy() for o in outs]
    trials = float(n) * nq * nw * nl
    fpt = gtls_trial_flops(nw)
    ach = fpt * trials / (kms * 1e-3) / 1e12
    pos = (fast_out[0] > 0) & (par_out[0] > 0)
    rel = np.abs(fast_out[0][pos] - par_out[0][pos]) / fast_out[0][pos]
    log(f"gtls fast {kms:.2f} ms ({trials / kms * 1e3:.3e} trials/s), parity {pms:.2f} ms; "
        f"fast vs float32 parity S/N rel. diff median {np.median(rel):.2e}")
    # CPU: the float32 restatement over bounded slices of epochs
    from oracle import timing
    cores, model = cpu_info()
    r = timing.gtls_rates(lc, tu, nw, nl, nsc, wmax, dw, deltat, nk=300)
    log(f"cpu gtls: 1 core {r['one_core']:.1f}, {r['cores']} processes {r['all_cores']:.1f} "
        f"epochs/s")
    return dict(metric="light curves/s and (epoch, W, L) trials/s (gtls single-transit search)",
                value=nq / (kms * 1e-3), unit="light curves/s", ms_per_step=wall * 1e3,
                dtype="f64", trials_per_s=trials / (kms * 1e-3),
                parity=dict(value=nq / (pms * 1e-3), unit="light curves/s", kernel_ms=pms,
                            dtype="f32", sn_rel_diff_median=float(np.median(rel)),
                            sn_rel_diff_p99=float(np.percentile(rel, 99)),
                            note="the reference's float32 sequential scans, bit-exact mode"),
                config=dict(workload=f"gtls normal mode, {nq} curves x {n} samples at 0.0204 d "
                                     f"(2% masked), nw {nw} widths 1-6 d, nl {nl}, nsc {nsc}",
                            n=n, nq=nq, nw=nw, nl=nl, nsc=nsc),
                roofline=dict(bound="fp64-valu", achieved=ach, peak=FP64_PEAK, unit="TFLOP/s",
                              frac=ach / FP64_PEAK, traffic=None, kernel="gtls_fast_kernel",
                              kernel_ms=kms, flops_per_trial=fpt),
                cpu_baseline=dict(value=r["all_cores"] / n, unit="light curves/s",
                                  cores=r["cores"], kind="port", one_core=r["one_core"] / n,
                                  sample=f"300 epochs per process of one curve with the float32 "
                                         f"C restatement, scaled to {n} epochs ({model})"))


# ----------------------------------------------------------------------------
# tlsfit workload: the reference's precise refit (works/sp/gtls_tesslc.py:
# 299-314: 10 t0 x 10 W x 20 L around a candidate) on a 27.4 d TESS sector at
# 2 min cadence, through gtrap_amd.tls.tlsfit; and a 1e6-trial scan of the
# same curve for the kernel's rate
# ----------------------------------------------------------------------------
# float64 operations per visited window sample of one trial in tlsgrid_kernel:
# mean pass 3 (offset, sum, count), region pass 5 (offset, centring and the
# region sums, B/D's time moments averaged in), residual pass 6
TLSGRID_FLOPS_PER_SAMPLE = 14
TLSFIT_REFITS = 100


def bench_tlsfit(steps: int, warmup: int):
    import torch
    from gtrap_amd import tls
    from oracle import tlsfold as ot
    rng = np.random.default_rng(1009)
    n = int(27.4 * 720)
    t = np.arange(n) * (2.0 / 1440.0)
    x = 1.0 + 5e-4 * rng.standard_normal(n) + tls.gen_trapzoid(t, -2e-3, 0.3, 0.05, 13.7)
    x[rng.random(n) < 0.01] = np.nan
    cands = rng.uniform(1.0, 26.4, TLSFIT_REFITS)

    def grids(T0p, Wp=0.3):
        tl = np.linspace(T0p - 0.1, T0p + 0.1, 10)
        Wl = np.linspace(Wp - Wp / 100, Wp + Wp / 100, 10)
        Ll = np.linspace(0.0, Wp / 2.0, 20)
        return Wl, Ll, tl

    for _ in range(max(1, warmup)):
        tls.tlsfit(t, x, *grids(cands[0]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in cands:
            tls.tlsfit(t, x, *grids(c))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    refit_ms = wall / (steps * TLSFIT_REFITS) * 1e3
    # kernel rate: 10 W x 20 L x 5000 t0 over the curve
    Wl, Ll = np.linspace(0.1, 0.5, 10), np.linspace(0.0, 0.2, 20)
    tl = np.linspace(0.2, 27.2, 5000)
    tls.grid_device(t, x, Wl, Ll, tl)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    tls.grid_device(t, x, Wl, Ll, tl)
    b.record()
    torch.cuda.synchronize()
    kms = a.elapsed_time(b)
    ntr = len(Wl) * len(Ll) * len(tl)
    win = sum(len(Ll) * float(np.sum(np.searchsorted(t, tl + w / 2, "right")
                                     - np.searchsorted(t, tl - w / 2, "left"))) for w in Wl)
    ach = TLSGRID_FLOPS_PER_SAMPLE * win / (kms * 1e-3) / 1e12
    log(f"tlsfit refit {refit_ms:.3f} ms; scan kernel {kms:.3f} ms ({ntr} trials)")
    # CPU: the numpy restatement of get_chi2Hmax over 15 refits (about 10 s)
    cores, model = cpu_info()
    ncpu = 15
    c0 = time.perf_counter()
    for c in cands[:ncpu]:
        Wc, Lc, tc = grids(c)
        for w in Wc:
            for l in Lc:
                for k in tc:
                    ot.chi2hmax_numpy(t - k, x, w, l)
    csec = (time.perf_counter() - c0) / ncpu
    return dict(metric="candidate refits/s (tls.py tlsfit, 2000 trials each) and trials/s",
                value=1e3 / refit_ms, unit="refits/s", ms_per_step=wall / steps * 1e3,
                dtype="f64", refit_ms=refit_ms, scan_trials_per_s=ntr / (kms * 1e-3),
                config=dict(workload=f"tlsfit refits of 10 t0 x 10 W x 20 L on a 27.4 d sector "
                                     f"at 2 min ({n} samples, 1% NaN), {TLSFIT_REFITS} "
                                     f"candidates per step; scan 10 W x 20 L x 5000 t0",
                            n=n, refits_per_step=TLSFIT_REFITS),
                roofline=dict(bound="fp64", achieved=ach, peak=FP64_PEAK, unit="TFLOP/s",
                              frac=ach / FP64_PEAK, traffic=None, kernel="tlsgrid_kernel (scan)",
                              kernel_ms=kms, window_samples_per_trial=win / ntr,
                              flops_per_window_sample=TLSGRID_FLOPS_PER_SAMPLE),
                cpu_baseline=dict(value=1.0 / csec, unit="refits/s", cores=1, kind="port",
                                  sample=f"{ncpu} refits of 2000 trials with the numpy "
                                         f"restatement of get_chi2Hmax ({model})"))


WORKLOADS = {"gfilter": bench_gfilter, "geebls": bench_geebls, "serve": bench_serve,
             "copy": bench_copy, "precision": bench_precision, "gtls": bench_gtls,
             "tlsfit": bench_tlsfit, "tls1": bench_tls1, "openloop": bench_openloop}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="serve", choices=sorted(WORKLOADS))
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the transit search shards by curve with no exchange between GPUs)",
              file=sys.stderr)
        return 2
    import torch
    from gtrap_amd import _native
    _native.require_gpu()
    torch.cuda.init()
    res = WORKLOADS[args.workload](args.steps, args.warmup)
    line = dict(metric=res.pop("metric"), value=res.pop("value"), unit=res.pop("unit"),
                n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=res.pop("ms_per_step"),
                higher_is_better=res.pop("higher_is_better", True), scaling="weak",
                vs_baseline=None, dtype=res.pop("dtype"), data="synthetic", **res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
