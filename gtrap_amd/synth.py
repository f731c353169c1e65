"""Seeded synthetic light curves for the benchmark and test workloads.

Generation is in float64 with ``numpy.random.default_rng(seed)`` and the data
are stored float32, as the workloads prescribe.  Every generator returns the
interleaved (n, nq) layout of the reference's image (element (i, q) at
``i*nq + q``; gtrap/read_keplerlc.py:49-50).
"""
from __future__ import annotations

import numpy as np

CADENCE = 0.0204  # days (Kepler long cadence, the reference's deltat scale)
INVAL = -1000.0   # invalid-time marker (gtrap/read_keplerlc.py:27)


def trapezoid_dip(t, t0, period, depth, t14, ingress):
    """Periodic trapezoid dip of total duration t14 and ingress/egress ``ingress``.

    Returns a non-negative dip profile (0 outside transit, ``depth`` at the
    flat bottom).  ``period=None`` gives a single event at t0.
    """
    t = np.asarray(t, dtype=np.float64)
    if period is None:
        dtc = t - t0
    else:
        dtc = (t - t0 + 0.5 * period) % period - 0.5 * period
    a = np.abs(dtc)
    half = 0.5 * t14
    flat = half - ingress
    prof = np.zeros_like(t)
    inside = a < half
    prof[inside] = depth
    edge = inside & (a > flat)
    prof[edge] = depth * (half - a[edge]) / max(ingress, 1e-12)
    return prof


def box_dip(t, t0, period, depth, duration):
    dtc = (np.asarray(t, np.float64) - t0 + 0.5 * period) % period - 0.5 * period
    return np.where(np.abs(dtc) < 0.5 * duration, depth, 0.0)


def config1(seed: int = 1001):
    """One curve, 4096 samples at 0.0204 d, sigma 1e-3, one periodic trapezoid
    (depth 2e-3, P 3.7 d, T14 0.12 d, ingress 0.01 d, random epoch)."""
    rng = np.random.default_rng(seed)
    n = 4096
    t = np.arange(n) * CADENCE
    P, t14, ing, depth = 3.7, 0.12, 0.01, 2e-3
    epoch = rng.uniform(0, P)
    flux = 1.0 + rng.normal(0, 1e-3, n) - trapezoid_dip(t, epoch, P, depth, t14, ing)
    periods = np.geomspace(0.5, 20.0, 1000)
    widths = np.linspace(0.04, 0.3, 10)
    return dict(t=t.astype(np.float32), flux=flux.astype(np.float32), periods=periods,
                widths=widths, nl=10, truth=dict(P=P, epoch=epoch, t14=t14, ingress=ing,
                                                  depth=depth))


def config2(nq: int = 1024, n: int = 16384, seed: int = 1002):
    """Batched geebls workload: non-common time (1% cadence jitter per curve),
    5% masked in 3 gaps (time set to INVAL, flux zig-zag filled with the
    curve's min/max as gtrap/read_keplerlc.py:39-46), AR(1) red noise
    (phi 0.9, sigma 5e-4) plus white noise 1e-3, boxes in half the curves."""
    rng = np.random.default_rng(seed)
    lc = np.empty((n, nq), np.float32)
    tu = np.empty((n, nq), np.float32)
    ntrue = np.empty(nq, np.uint32)
    truth = []
    nmask = int(round(0.05 * n))
    for q in range(nq):
        dt = CADENCE * (1.0 + rng.uniform(-0.01, 0.01))
        t = np.arange(n) * dt
        e = rng.normal(0, 5e-4, n)
        red = np.empty(n)
        acc = 0.0
        for i in range(n):  # AR(1)
            acc = 0.9 * acc + e[i]
            red[i] = acc
        f = 1.0 + red + rng.normal(0, 1e-3, n)
        P = None
        if rng.random() < 0.5:
            P = rng.uniform(1.0, 30.0)
            depth = rng.uniform(5e-4, 5e-3)
            qdur = rng.uniform(11, 22) / 1024
            f -= box_dip(t, rng.uniform(0, P), P, depth, qdur * P)
        truth.append(P)
        # three non-overlapping gaps totalling nmask samples
        cuts = np.sort(rng.choice(np.arange(1, nmask), 2, replace=False))
        lens = np.diff(np.concatenate([[0], cuts, [nmask]]))
        starts = np.sort(rng.choice(np.arange(0, n - nmask - 3), 3, replace=False))
        starts = starts + np.arange(3) * 0  # positions in the gap-free index space
        mask = np.zeros(n, bool)
        pos = 0
        free = np.arange(n)
        offs = 0
        for s0, L in zip(starts, lens):
            a = s0 + offs
            mask[a:a + L] = True
            offs += L
        tt = t.copy()
        tt[mask] = INVAL
        ff = f.copy()
        lo, hi = ff[~mask].min(), ff[~mask].max()
        m2 = mask.copy()
        ev = np.zeros(n, bool); ev[::2] = True
        ff[m2 & ev] = lo
        ff[m2 & ~ev] = hi
        lc[:, q] = ff
        tu[:, q] = tt
        ntrue[q] = int((~mask).sum())
    return dict(lc=lc, tu=tu, ntrue=ntrue, truth=truth)


def config3(nq: int = 4096, n: int = 65536, seed: int = 1003, device=None):
    """gfilter workload: 1 + quadratic trend + sinusoid + noise 1e-3 + 0.1%
    outliers of +-1e-2.  With ``device`` the image is generated with torch on
    that device (same distribution, torch's generator)."""
    if device is not None:
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        dev = torch.device(device)
        tau = (torch.arange(n, device=dev, dtype=torch.float64) / n)[:, None]
        t = (torch.arange(n, device=dev, dtype=torch.float64) * CADENCE)[:, None]
        u = lambda lo, hi, shape: lo + (hi - lo) * torch.rand(shape, generator=g, device=dev,
                                                              dtype=torch.float64)
        c1, c2 = u(-1e-2, 1e-2, (1, nq)), u(-1e-2, 1e-2, (1, nq))
        ps, ph = u(5, 50, (1, nq)), u(0, 2 * np.pi, (1, nq))
        img = torch.empty((n, nq), dtype=torch.float32, device=dev)
        chunk = max(1, (1 << 24) // nq)
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            f = 1.0 + c1 * tau[a:b] + c2 * tau[a:b] ** 2 + 1e-3 * torch.sin(
                2 * np.pi * t[a:b] / ps + ph)
            f = f + 1e-3 * torch.randn((b - a, nq), generator=g, device=dev, dtype=torch.float64)
            out = torch.rand((b - a, nq), generator=g, device=dev) < 1e-3
            sgn = torch.where(torch.rand((b - a, nq), generator=g, device=dev) < 0.5, -1.0, 1.0)
            f = torch.where(out, f + 1e-2 * sgn, f)
            img[a:b] = f.to(torch.float32)
        return img
    rng = np.random.default_rng(seed)
    tau = (np.arange(n) / n)[:, None]
    t = (np.arange(n) * CADENCE)[:, None]
    c1, c2 = rng.uniform(-1e-2, 1e-2, (2, 1, nq))
    ps = rng.uniform(5, 50, (1, nq))
    ph = rng.uniform(0, 2 * np.pi, (1, nq))
    f = 1.0 + c1 * tau + c2 * tau ** 2 + 1e-3 * np.sin(2 * np.pi * t / ps + ph)
    f += rng.normal(0, 1e-3, (n, nq))
    out = rng.random((n, nq)) < 1e-3
    f[out] += 1e-2 * np.where(rng.random(out.sum()) < 0.5, -1.0, 1.0)
    return f.astype(np.float32)
