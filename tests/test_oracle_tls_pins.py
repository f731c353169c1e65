"""Pins of the trapezoid statistic's residual and S/N leg, and of the
periodogram statistics, by what the mathematics fixes rather than by a second
transcription of gtrap/tls.py:44-79.

* Residual identity: tls.py's residual (:74-77) is the sum of squares of
  x - mean - model over the window, where the model is the trapezoid of
  gen_trapzoid (tls.py:2-15) with the fitted H.  gen_trapzoid is continuous, so
  the identity does not depend on which region a boundary sample falls in.
* Noiseless known answers: on an exact trapezoid the closed form recovers H
  (tls.py:113-122; also on random continuous times) and the residual is 0.
* tlsstats: a hand-computed closed form.
"""
import numpy as np

from oracle import tlsfold as ot


def gen_trapzoid(t, H, W, L, t0):
    """tls.py:2-15 restated: -H/2 outside, H/2 on the flat bottom, ramps of
    width L centred at -W/4 and W/4."""
    tn = np.asarray(t, np.float64) - t0
    x = np.zeros(len(tn))
    x[tn <= -W / 4 - L / 2] = -H / 2
    x[tn >= W / 4 + L / 2] = -H / 2
    x[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = H / 2
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    x[mB] = H / L * (tn[mB] + W / 4)
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    x[mD] = -H / L * (tn[mD] - W / 4)
    return x


def _direct_residual(t, x, P, e, W, L, dtep, H, divx=True):
    """sum over the window [-W/2, W/2) of (x - mean - gen_trapzoid(d, H))^2,
    mean over the closed window (tls.py:57-59)."""
    d, ok = ot.fold_offsets(t, P, e, dtep, x)
    xx = np.asarray(x, np.float64)[ok]
    mean = xx[(d >= -W / 2) & (d <= W / 2)].mean() if divx else 0.0
    win = (d >= -W / 2) & (d < W / 2)
    r = xx[win] - mean - gen_trapzoid(d[win], H, W, L, 0.0)
    return float(np.sum(r * r)), int(win.sum()), mean


def test_model_is_continuous_at_the_region_boundaries():
    """The identity below needs gen_trapzoid to be continuous: its value at each
    boundary equals both one-sided limits."""
    H, W, L = -2.0, 1.0, 0.2
    for b in (-W / 4 - L / 2, -W / 4 + L / 2, W / 4 - L / 2, W / 4 + L / 2):
        v = gen_trapzoid(np.array([b - 1e-12, b, b + 1e-12]), H, W, L, 0.0)
        np.testing.assert_allclose(v, v[1], atol=1e-10)


def test_kat_tls_py_residual_is_zero():
    """tls.py:113-122: H = -3, W = 30, L = 5 at t0 = 50 on linspace(0, 100, 132)
    without mean removal: H is recovered and the residual vanishes."""
    t = np.linspace(0, 100, 132)
    x = gen_trapzoid(t, -3.0, 30.0, 5.0, 50.0)
    # a period far above the baseline makes the fold the identity; W = 30,
    # nl = 27, m = 10 give L = 5
    sn, H, n, res, mean = ot.trial_res(t.astype(np.float32), x.astype(np.float32), 1e6, 50, 30.0,
                                       10, 27, 1.0, divx=False)
    assert abs(H + 3.0) < 1e-5 and mean == 0.0
    # float32 storage of t and x: the model misses x by ~1e-7 per sample
    assert res < 1e-9 * float(np.sum(x * x))
    assert sn > 1e3


def test_kat_random_times_recovers_h_exactly():
    """The closed form Hmax = -numB/denA is exact for a noiseless trapezoid on
    any sampling (tls.py:71-73): random continuous times, random shapes."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        t = np.sort(rng.uniform(0, 10, 400))
        W = rng.uniform(0.5, 3.0)
        nl = 10
        m = int(rng.integers(1, nl))
        L = m * (0.45 * W / nl)
        t0 = float(rng.uniform(3, 7))
        H = float(rng.uniform(-1, 1))
        x = gen_trapzoid(t, H, W, L, t0)
        # float64 time and flux through the numpy restatement (exact inputs)
        Hn, res, n, _ = ot.chi2hmax_numpy(t - t0, x, W, L, divx=False)
        # denA and numB cancel between regions: float64 rounding, 1e-9
        assert abs(Hn - H) <= 1e-9 * abs(H)
        assert res <= 1e-16 * H * H * n


def test_kat_aligned_grid_with_mean_removal():
    """With samples at the midpoints of a grid aligned to the region boundaries
    the model has zero mean over the window, so mean removal (divx=True, the
    setting of the periodic search) leaves the trapezoid intact: H exact,
    residual 0.  Dyadic values keep t and x exact in float32."""
    dt = 1.0 / 64
    W, nl, m = 2.0, 9, 5
    L = m * (0.45 * W / nl)          # 0.5 = 32 dt; W/4 - L/2 = 0.25 = 16 dt
    assert abs(L - 0.5) < 1e-15
    e = 400
    t0 = e * dt
    k = np.arange(-int(W / 2 / dt), int(W / 2 / dt))
    t = t0 + (k + 0.5) * dt
    H = -2.0 ** -6
    x = gen_trapzoid(t, H, W, L, t0)
    assert abs(x.mean()) < 1e-18
    t32, x32 = t.astype(np.float32), x.astype(np.float32)
    assert np.array_equal(t32.astype(np.float64), t) and np.array_equal(x32.astype(np.float64), x)
    sn, Hc, n, res, mean = ot.trial_res(t32, x32, 1e6, e, W, m, nl, dt, divx=True)
    assert n == len(t) and abs(mean) < 1e-18
    assert abs(Hc - H) <= 1e-15 and res <= 1e-28
    assert sn == np.inf or sn > 1e6


def test_residual_identity_random_folds():
    """C oracle residual == direct residual against gen_trapzoid with the
    fitted H, on random continuous times folded at random periods."""
    rng = np.random.default_rng(12)
    n = 900
    t = np.sort(rng.uniform(0, 30, n)).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    t[rng.random(n) < 0.03] = -1000.0
    x[rng.random(n) < 0.01] = np.nan
    dtep = 0.0204
    checked = 0
    for _ in range(300):
        P = float(rng.uniform(0.6, 12.0))
        e = int(rng.integers(0, int(np.ceil(P / dtep))))
        W = float(rng.uniform(0.04, 0.5))
        nl = 10
        m = int(rng.integers(1, nl))
        L = m * (0.45 * W / nl)
        sn, H, cnt, res, mean = ot.trial_res(t, x, P, e, W, m, nl, dtep)
        rd, nd, md = _direct_residual(t, x, P, e, W, L, dtep, H)
        assert cnt == nd
        if nd <= 3:
            assert np.isnan(sn)
            continue
        np.testing.assert_allclose(mean, md, rtol=1e-12)
        np.testing.assert_allclose(res, rd, rtol=1e-9, atol=1e-18)
        np.testing.assert_allclose(sn, -H / np.sqrt(rd / (nd - 3)), rtol=1e-9)
        checked += 1
    assert checked > 200


def test_nonfinite_flux_is_skipped_like_a_masked_sample():
    rng = np.random.default_rng(13)
    n = 700
    t = (np.arange(n) * 0.0204).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    bad = rng.choice(n, 12, replace=False)
    x2 = x.copy()
    x2[bad[:6]] = np.nan
    x2[bad[6:9]] = np.inf
    x2[bad[9:]] = -np.inf
    t2 = t.copy()
    t2[bad] = -1000.0
    per = np.geomspace(0.5, 4.0, 9)
    widths = np.linspace(0.04, 0.3, 10)
    a = ot.tlsfold(t, x2, per, widths, 10, 0.0204)
    b = ot.tlsfold(t2, x, per, widths, 10, 0.0204)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_tlsstats_closed_form():
    sn = np.array([1.0, 3.0, np.nan, -np.inf, 3.0, 2.0])
    best, val, sde = ot.tlsstats(sn)
    # finite values 1, 3, 3, 2: mean 9/4, variance 23/4 - 81/16 = 11/16
    assert best == 1 and val == 3.0
    np.testing.assert_allclose(sde, 0.75 / np.sqrt(11.0 / 16.0), rtol=1e-15)
    # a perfect fit (+inf) wins but stays out of the mean and variance
    best, val, sde = ot.tlsstats(np.array([1.0, np.inf, 2.0, np.inf]))
    assert best == 1 and val == np.inf and sde == np.inf
    # no trial anywhere
    best, val, sde = ot.tlsstats(np.array([np.nan, -np.inf]))
    assert best == -1 and val == -np.inf and np.isnan(sde)


def test_tlsstats_equals_getstat_formula_on_finite_series():
    """On a finite series the SDE is getstat.py:114-118's (max - ave)/std with
    ave = sum/n, var = sumsq/n - ave^2, restated by the BLS statistics oracle."""
    from oracle import getstat as og
    rng = np.random.default_rng(14)
    v = rng.uniform(1, 5, 2048).astype(np.float32)
    best, val, sde = ot.tlsstats(v.astype(np.float64))
    pest, sde_b, _, _, idx = og.blsstat(v, None, None, 2048, 1, 1.0, 1.0)
    assert best == int(idx[0])
    np.testing.assert_allclose(sde, sde_b[0], rtol=1e-12)


def test_detection_statistic_closed_form():
    """stat="dchi2": the least-squares fit of the trapezoid shape
    g = gen_trapzoid(d, 1, W, L, 0) to y = x - mean over the window, computed
    here directly with numpy: H* = b/a, D = sign(-H*) sqrt(b^2/a) / sigma with
    sigma the standard deviation of the curve's usable fluxes."""
    rng = np.random.default_rng(15)
    n = 900
    t = np.sort(rng.uniform(0, 30, n)).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    t[rng.random(n) < 0.03] = -1000.0
    sigma = float(np.std(x[t >= 0].astype(np.float64), ddof=1))
    dtep = 0.0204
    checked = 0
    for _ in range(200):
        P = float(rng.uniform(0.6, 12.0))
        e = int(rng.integers(0, int(np.ceil(P / dtep))))
        W = float(rng.uniform(0.04, 0.5))
        nl, m = 10, int(rng.integers(1, 10))
        L = m * (0.45 * W / nl)
        D, Hs, cnt = ot.trial(t, x, P, e, W, m, nl, dtep, stat="dchi2")
        d, ok = ot.fold_offsets(t, P, e, dtep, x)
        xx = x[ok].astype(np.float64)
        mean = xx[(d >= -W / 2) & (d <= W / 2)].mean()
        win = (d >= -W / 2) & (d < W / 2)
        y = xx[win] - mean
        g = gen_trapzoid(d[win], 1.0, W, L, 0.0)
        a, b = float(g @ g), float(y @ g)
        if win.sum() <= 3:
            assert np.isnan(D)
            continue
        want = -np.sign(b) * np.sqrt(b * b / a) / sigma
        # the oracle's sigma is the sum-of-squares form (flux near 1: ~1e-9)
        np.testing.assert_allclose(D, want, rtol=1e-8)
        np.testing.assert_allclose(Hs, b / a, rtol=1e-9)
        checked += 1
    assert checked > 150
    # L = 0 is no trial in both statistics
    assert np.isnan(ot.trial(t, x, 3.0, 10, 0.2, 0, 10, dtep, stat="dchi2")[0])


def test_detection_statistic_recovers_the_single_curve_period():
    """The single-curve workload (trapezoid of depth 2e-3 at P = 3.7 d in
    noise 1e-3): on every tenth period of the grid plus 3.7 d, the detection
    statistic's best period is the injected one with H* near the depth;
    tls.py's S/N picks another period (its value does not grow with the
    number of transits in a window)."""
    from gtrap_amd import synth
    d = synth.tls_single_curve()
    per = np.sort(np.concatenate([d["periods"][::10], [3.7]]))
    r = ot.tlsfold(d["t"], d["flux"], per, d["widths"], 10, 0.0204, stat="dchi2")
    best, val, sde = ot.tlsstats(r["sn"])
    assert per[best] == 3.7 and val > 10 and sde > 5
    assert abs(r["h"][best] + 2e-3) < 5e-4
    r0 = ot.tlsfold(d["t"], d["flux"], per, d["widths"], 10, 0.0204)
    assert per[ot.tlsstats(r0["sn"])[0]] != 3.7
