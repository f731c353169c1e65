"""Timing of the CPU restatements for the benchmark's CPU baseline (test
infrastructure: only bench.py's cpu_baseline leg uses this, as the thing it
reports beside the GPU path, never as part of the product)."""
from __future__ import annotations

import os
import time

import numpy as np


def periods_worker(kind, t, x, periods, widths, nl, dtep, cap):
    """Search ``periods`` of one curve on one core, in the given order, until
    ``cap`` seconds are spent, with the C oracle (kind "c") or the numpy
    restatement of tls.py (kind "numpy"); returns (periods done, trial
    evaluations done, seconds)."""
    for v in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS"):
        os.environ[v] = "1"
    from . import tlsfold as ot
    t1 = time.perf_counter()
    trials = done = 0
    for P in periods:
        if kind == "c":
            ot.tlsfold(t, x, np.array([P]), widths, nl, dtep)
        else:
            ot.period_numpy(t, x, P, widths, nl, dtep)
        trials += int(np.ceil(P / dtep)) * len(widths) * nl
        done += 1
        if time.perf_counter() - t1 > cap:
            break
    return done, trials, time.perf_counter() - t1


def trial_rates(t, x, periods, widths, nl, dtep, cap=8.0, max_procs=16, log=None, seed=0):
    """CPU rates of the trapezoid search over one curve's period grid, for the
    C oracle and the numpy tls.py restatement, on one core and on a process
    pool over the host cores (at most ``max_procs``).  Each leg runs about
    ``cap`` seconds on periods drawn at random from the grid (a capped run
    then covers short and long periods in proportion), and reports trial
    evaluations/s and whole curves/s = periods/s over the grid's size."""
    import concurrent.futures as cf
    import multiprocessing as mp
    cores = max(1, min(max_procs, len(os.sched_getaffinity(0))))
    out = dict(cores=cores)
    m = len(periods)
    periods = np.asarray(periods, np.float64)
    rng = np.random.default_rng(seed)
    for kind in ("c", "numpy"):
        order = rng.permutation(m)
        pd, tr, sec = periods_worker(kind, t, x, periods[order], widths, nl, dtep, cap)
        out[f"{kind}_1core"] = tr / sec
        out[f"{kind}_1core_curves"] = pd / sec / m
        shares = [periods[rng.permutation(m)] for _ in range(cores)]
        t1 = time.perf_counter()
        with cf.ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
            futs = [ex.submit(periods_worker, kind, t, x, sh, widths, nl, dtep, cap)
                    for sh in shares]
            done = [f.result(timeout=20 * cap + 60) for f in futs]
        wall = time.perf_counter() - t1
        out[f"{kind}_allcores"] = sum(d[1] for d in done) / wall
        out[f"{kind}_allcores_curves"] = sum(d[0] for d in done) / wall / m
        if log:
            log(f"cpu {kind}: 1 core {out[f'{kind}_1core']:.3e}, {cores} processes "
                f"{out[f'{kind}_allcores']:.3e} trial evals/s")
    return out


def gtls_epochs_worker(lc, tu, nw, nl, nsc, wmax, dw, deltat, k0, k1):
    """The float32 gtls restatement over epochs [k0, k1) of the given curves;
    returns seconds."""
    for v in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS"):
        os.environ[v] = "1"
    from . import gtls as og
    t1 = time.perf_counter()
    og.gtls(lc, tu, nw, nl, nsc, wmax, dw, deltat, ks=(k0, k1))
    return time.perf_counter() - t1


def gtls_rates(lc, tu, nw, nl, nsc, wmax, dw, deltat, nk=400, max_procs=16):
    """Epochs/s of the float32 gtls restatement on one core and on a process
    pool (each process a different slice of nk epochs of curve 0)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    cores = max(1, min(max_procs, len(os.sched_getaffinity(0))))
    nt = lc.shape[0]
    c1 = lc[:, :1].copy()
    t1_ = tu[:, :1].copy()
    k0 = nt // 2
    sec = gtls_epochs_worker(c1, t1_, nw, nl, nsc, wmax, dw, deltat, k0, k0 + nk)
    starts = [int(v) for v in np.linspace(nsc, nt - nsc - nk, cores)]
    t0 = time.perf_counter()
    with cf.ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
        futs = [ex.submit(gtls_epochs_worker, c1, t1_, nw, nl, nsc, wmax, dw, deltat, a, a + nk)
                for a in starts]
        for f in futs:
            f.result(timeout=600)
    wall = time.perf_counter() - t0
    return dict(cores=cores, one_core=nk / sec, all_cores=cores * nk / wall)
