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


def tls_single_curve(seed: int = 1001):
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


def bls_batch(nq: int = 1024, n: int = 16384, seed: int = 1002):
    """Batched geebls workload: non-common time (1% cadence jitter per curve),
    5% of samples masked in 3 gaps (time set to INVAL, flux zig-zag filled
    with the curve's min/max as gtrap/read_keplerlc.py:39-46), AR(1) red noise
    (phi 0.9, innovation sigma 5e-4) plus white noise 1e-3, box transits
    (depth U(5e-4, 5e-3), P U(1, 30) d, duration fraction U(11, 22)/1024 of P)
    in a random half of the curves.  Returns lc, tu (n, nq), ntrue, truth."""
    from scipy.signal import lfilter
    rng = np.random.default_rng(seed)
    lc = np.empty((n, nq), np.float32)
    tu = np.empty((n, nq), np.float32)
    ntrue = np.empty(nq, np.uint32)
    truth = []
    nmask = int(round(0.05 * n))
    even = (np.arange(n) % 2) == 0
    for q in range(nq):
        dt = CADENCE * (1.0 + rng.uniform(-0.01, 0.01))
        t = np.arange(n) * dt
        red = lfilter([1.0], [1.0, -0.9], rng.normal(0, 5e-4, n))
        f = 1.0 + red + rng.normal(0, 1e-3, n)
        item = None
        if rng.random() < 0.5:
            P = rng.uniform(1.0, 30.0)
            item = dict(P=P, depth=rng.uniform(5e-4, 5e-3), q=rng.uniform(11, 22) / 1024,
                        epoch=rng.uniform(0, P))
            f -= box_dip(t, item["epoch"], P, item["depth"], item["q"] * P)
        truth.append(item)
        # three non-overlapping gaps whose lengths sum to nmask
        cuts = np.sort(rng.choice(np.arange(1, nmask), 2, replace=False))
        lens = np.diff(np.concatenate([[0], cuts, [nmask]]))
        slots = np.sort(rng.choice(n - nmask, 3, replace=False))
        mask = np.zeros(n, bool)
        shift = 0
        for s0, L in zip(slots, lens):
            mask[s0 + shift:s0 + shift + L] = True
            shift += L
        tt = np.where(mask, INVAL, t)
        lo, hi = f[~mask].min(), f[~mask].max()
        f = np.where(mask & even, lo, np.where(mask & ~even, hi, f))
        lc[:, q] = f
        tu[:, q] = tt
        ntrue[q] = int((~mask).sum())
    return dict(lc=lc, tu=tu, ntrue=ntrue, truth=truth)


def bls_batch_grid(n: int = 16384):
    """Trial grid of the batched geebls workload: fmin = 1/30, nf = 2048
    frequencies up to 1 d^-1, nb = 1024 bins, widths kmi < k <= kma with
    kmi = 10, kma = 22 (12 widths), kkmi = int(n*kmi/nb), block 512."""
    nf, nb = 2048, 1024
    fmin = 1.0 / 30.0
    df = (1.0 - fmin) / (nf - 1)
    kmi, kma = 10, 22
    return dict(nf=nf, nb=nb, kma=kma, kmi=kmi, kkmi=int(n * kmi / nb), fmin=fmin, df=df,
                nthread=512)


def detrend_image(nq: int = 4096, n: int = 65536, seed: int = 1003, device=None):
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


def precision_batch(nq: int = 256, n: int = 262144, seed: int = 1005, period: float = 7.3,
            nf: int = 128):
    """Precision workload: nq curves of n samples at a 2-minute cadence (364 d
    for n = 262144), one periodic trapezoid per curve at P = ``period``
    (random epoch, duty cycle q ~ U(11, 22)/1024, ingress T14/10, depth 1e-4),
    white noise with sigma set per curve so that depth*sqrt(n q)/sigma is
    U(7, 15).  The BLS grid is nf frequencies centred on 1/P with step
    0.25 * (11/1024) / baseline (every transit resolved), nb = 1024,
    kmi = 10, kma = 23, kkmi = n*kmi/nb, nthread 256."""
    rng = np.random.default_rng(seed)
    cad = 2.0 / 1440.0
    t = np.arange(n) * cad
    base = n * cad
    lc = np.empty((n, nq), np.float32)
    snr = rng.uniform(7, 15, nq)
    qd = rng.uniform(11, 22, nq) / 1024
    depth = 1e-4
    for k in range(nq):
        sigma = depth * np.sqrt(n * qd[k]) / snr[k]
        t14 = qd[k] * period
        f = 1.0 + rng.normal(0, sigma, n)
        f -= trapezoid_dip(t, rng.uniform(0, period), period, depth, t14, t14 / 10)
        lc[:, k] = f
    nb, kmi, kma = 1024, 10, 23
    df = 0.25 * (11 / 1024) / base
    fmin = 1.0 / period - 0.5 * nf * df
    return dict(t64=t, lc=lc, offset=lc.astype(np.float64).mean(0).astype(np.float32),
                ntrue=np.full(nq, n, np.uint32),
                grid=dict(nf=nf, nb=nb, kmi=kmi, kma=kma, kkmi=int(n * kmi / nb), fmin=fmin,
                          df=df, nthread=256),
                truth=dict(P=period, snr=snr, q=qd, depth=depth))


def serving_trace(count: int = 200, seed: int = 1004, rate: float = 2000.0,
                     nlo: int = 2048, nhi: int = 70000, period_stride: int = 1):
    """Serving trace: Poisson arrivals (mean ``rate`` req/s), lengths
    log-uniform nlo..nhi at 0.0204 d cadence, flux = 1 + slow trend + noise
    1e-3, trapezoid transits (depth U(1e-3, 5e-3), P U(0.5, baseline/3),
    T14 U(0.05, 0.3), ingress T14/10) in 30% of the requests; periods
    geomspace(0.5, baseline/3, m) with m 500..4000 log-linear in n."""
    from .serve import Request
    from .tlsfold import period_grid
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate, count))
    out, truth = [], []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(nlo), np.log(nhi)))))
        t = np.arange(n) * CADENCE
        base = n * CADENCE
        flux = 1.0 + 2e-3 * np.sin(2 * np.pi * t / rng.uniform(5, 50) + rng.uniform(0, 6.3))
        flux += rng.normal(0, 1e-3, n)
        tr = None
        if rng.random() < 0.3:
            P = rng.uniform(0.5, base / 3)
            t14 = rng.uniform(0.05, 0.3)
            tr = dict(P=P, epoch=rng.uniform(0, P), depth=rng.uniform(1e-3, 5e-3), t14=t14)
            flux -= trapezoid_dip(t, tr["epoch"], P, tr["depth"], t14, t14 / 10)
        per = period_grid(base, n=n)[::period_stride]
        out.append(Request(i, t.astype(np.float32), flux.astype(np.float32), per,
                           arrival=float(arr[i])))
        truth.append(tr)
    return out, truth


def serving_stream(rate: float, duration: float, seed: int = 2004, nlo: int = 2048,
                   nhi: int = 70000):
    """Open-loop stream: Poisson arrivals at ``rate`` req/s over ``duration``
    seconds, the requests of :func:`serving_trace` (same length, flux, transit
    and period-grid distributions), generated in float32 to keep a long stream
    cheap.  Returns (requests, truths) sorted by arrival (seconds from 0)."""
    from .serve import Request
    from .tlsfold import period_grid
    rng = np.random.default_rng(seed)
    arr = []
    tt = 0.0
    while True:
        tt += rng.exponential(1.0 / rate)
        if tt >= duration:
            break
        arr.append(tt)
    out, truth = [], []
    for i, a in enumerate(arr):
        n = int(round(np.exp(rng.uniform(np.log(nlo), np.log(nhi)))))
        t = np.arange(n, dtype=np.float64) * CADENCE
        base = n * CADENCE
        flux = 1.0 + 2e-3 * np.sin(2 * np.pi * t / rng.uniform(5, 50) + rng.uniform(0, 6.3))
        flux += 1e-3 * rng.standard_normal(n, dtype=np.float32)
        tr = None
        if rng.random() < 0.3:
            P = rng.uniform(0.5, base / 3)
            t14 = rng.uniform(0.05, 0.3)
            tr = dict(P=P, epoch=rng.uniform(0, P), depth=rng.uniform(1e-3, 5e-3), t14=t14)
            flux -= trapezoid_dip(t, tr["epoch"], P, tr["depth"], t14, t14 / 10)
        out.append(Request(i, t.astype(np.float32), flux.astype(np.float32),
                           period_grid(base, n=n), arrival=float(a)))
        truth.append(tr)
    return out, truth


def recovered(best_period: float, truth, tol: float = 0.01,
              harmonics=(1.0 / 3, 0.5, 1.0, 2.0, 3.0)) -> bool:
    """Whether a reported period is within ``tol`` (relative) of the injected
    period or one of its harmonics."""
    if truth is None or not np.isfinite(best_period):
        return False
    return any(abs(best_period - k * truth["P"]) <= tol * k * truth["P"] for k in harmonics)


def trial_count(req, nw: int = 10, nl: int = 10, dtep: float = CADENCE) -> int:
    """(epoch, W, L) trial evaluations of one request."""
    return int(np.sum(np.ceil(np.asarray(req.periods) / dtep)) * nw * nl)
