"""Pins the gtls oracle (include/gtls_simple.h:1-406, include/gtls_absolute.h)."""
import numpy as np

from oracle import gtls as og
from oracle import tlsfold as ot


def _series(n=200, seed=0, dip=None):
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64) * 1.0
    x = 1e-2 * rng.standard_normal(n)
    if dip:
        k0, w, h = dip
        x[np.abs(t - k0) < w] += h
    return t.astype(np.float32), x.astype(np.float32)


def test_matches_tls_py_statistic_when_scoop_is_inside_window():
    """nsc < W/deltat: no sample precedes -W/2 (nSx = 0, the A.1 quirk is void),
    every scoop sample lies in [-W/2, W/2]; the float32 restatement then equals
    tls.py's get_chi2Hmax (float64) with the reference's sign (S/N = +H/...)."""
    t, x = _series(dip=(100, 2, 0.05))
    lc, tu = x[:, None], t[:, None]
    nsc, nl, wmax = 9, 5, 9.0  # window [-4.5, 4.5] holds exactly the 9 scoop samples
    out = og.gtls(lc, tu, 1, nl, nsc, wmax, 0.0, 1.0)
    for k in (50, 100, 150):
        best = 0.0
        for m in range(1, nl):
            sn, H, n = ot.trial_numpy(t, x, 1e9, k, wmax, m, nl, 1.0)
            best = max(best, -sn) if np.isfinite(sn) else best
        np.testing.assert_allclose(out["tlssn"][k, 0], best, rtol=2e-4)


def test_region_a_index_quirk_is_reproduced():
    """Region A sums the flux at scoop index i while testing the time at i+nSx
    (gtls_simple.h:214-221): a flux change before the window start moves S/N."""
    t, x = _series(seed=1, dip=(100, 3, 0.05))
    lc, tu = x[:, None].copy(), t[:, None]
    kw = dict(nw=4, nl=6, nsc=31, wmax=12.0, dw=1.0, deltat=1.0)
    base = og.gtls(lc, tu, **kw)["tlssn"][92:101, 0]
    lc2 = lc.copy()
    lc2[85, 0] += 0.3  # time 85: outside [-W/2, W/2] of every epoch 92..100
    moved = og.gtls(lc2, tu, **kw)["tlssn"][92:101, 0]
    assert np.any(moved != base)


def test_masked_samples_and_fill_factor():
    t, x = _series(seed=2, dip=(60, 2, 0.05))
    tu = t.copy()
    tu[55:66] = -1000.0  # the window around epoch 60 is mostly masked
    out = og.gtls(x[:, None], tu[:, None], 1, 5, 21, 10.0, 0.0, 1.0)
    # fill factor = valid/(W/deltat) < 0.75 -> no trial, S/N 0, W = wmax, L = H = 0
    assert out["tlssn"][60, 0] == 0.0 and out["tlsl"][60, 0] == 0.0
    assert out["tlsw"][60, 0] == np.float32(10.0)
    assert out["tlst0"][60, 0] == np.float32(60.0)


def test_absolute_mode_uses_abs_h():
    t, x = _series(seed=3, dip=(80, 3, -0.08))  # negative dip: normal mode S/N < 0
    kw = dict(nw=2, nl=5, nsc=31, wmax=10.0, dw=2.0, deltat=1.0)
    norm = og.gtls(x[:, None], t[:, None], **kw)
    ab = og.gtls(x[:, None], t[:, None], mode="absolute", **kw)
    assert ab["tlssn"][80, 0] > 0 and ab["tlshmax"][80, 0] < 0
    assert ab["tlssn"][80, 0] > norm["tlssn"][80, 0]


def _gtls_scoop(lc, tu, k, nsc, deltat):
    """Scoop of epoch k as include/gtls_simple.h:51-65,121-134 builds it:
    rows k - nsc/2 .. k - nsc/2 + nsc - 1, stored time -1000 + fl(tu + 1000)
    (float32), masked rows (stored time < 0 or outside the curve) dropped,
    shifted time tau = fl(T - deltat*k) with one rounding (a fused multiply-add)."""
    nt = len(lc)
    rows = np.arange(k - nsc // 2, k - nsc // 2 + nsc)
    rows = rows[(rows >= 0) & (rows < nt)]
    T = (np.float32(-1000.0) + (tu[rows].astype(np.float32) + np.float32(1000.0))).astype(
        np.float32)
    ok = T >= 0
    tau = (T[ok].astype(np.float64) - np.float64(np.float32(deltat)) * k).astype(np.float32)
    return tau.astype(np.float64), lc[rows][ok].astype(np.float64)


def test_residual_identity_of_the_best_trial():
    """The reported S/N = H/sqrt(res/(n-3)) (gtls_simple.h:359-369) implies a
    residual; it equals the sum over the window (-W/2, W/2] of
    (x - mean - gen_trapzoid(tau, H, W, L, 0))^2 with the reported H, W, L
    (mean over the closed window, gtls_simple.h:139-152).  The model of
    tls.py:2-15 is continuous, so region membership at the boundaries does not
    enter; float32 arithmetic bounds the agreement."""
    from tests.test_oracle_tls_pins import gen_trapzoid
    rng = np.random.default_rng(21)
    nt = 600
    t = (np.arange(nt) * 0.0204).astype(np.float32)
    x = (1e-3 * rng.standard_normal(nt)).astype(np.float32)
    x[(t > 5.0) & (t < 5.12)] += 3e-3
    tu = t.copy()
    tu[300:310] = -1000.0
    nw, nl, wmax, dw, deltat = 8, 6, 0.3, 0.03, np.float32(0.0204)
    nsc = int(wmax / deltat + 3)
    out = og.gtls(x[:, None], tu[:, None], nw, nl, nsc, wmax, dw, deltat)
    checked = 0
    for k in range(20, nt - 20, 7):
        sn = float(out["tlssn"][k, 0])
        if sn <= 0:
            continue
        W, L, H = (float(out[n][k, 0]) for n in ("tlsw", "tlsl", "tlshmax"))
        tau, xx = _gtls_scoop(x, tu, k, nsc, deltat)
        W32 = np.float32(W)
        lo, hi = np.float32(-0.5 * W32), np.float32(0.5 * W32)
        mean = xx[(tau >= lo) & (tau <= hi)].mean()
        win = (tau > lo) & (tau <= hi)
        r = xx[win] - mean - gen_trapzoid(tau[win], H, W, L, 0.0)
        res_direct = float(np.sum(r * r))
        res_implied = (win.sum() - 3) * (H / sn) ** 2
        np.testing.assert_allclose(res_implied, res_direct, rtol=2e-4)
        checked += 1
    assert checked > 30


def test_float64_gtls_matches_tls_py_when_quirks_are_void():
    """The float64 restatement equals tls.py's float64 statistic where the
    scoop lies inside the window (no region-A index shift); the reference's
    L is float32 (gtls_simple.h:69,167), so 1e-7."""
    t, x = _series(dip=(100, 2, 0.05))
    lc, tu = x[:, None], t[:, None]
    nsc, nl, wmax = 9, 5, 9.0
    out = og.gtls64(lc, tu, 1, nl, nsc, wmax, 0.0, 1.0)
    for k in (50, 100, 150):
        best = 0.0
        for m in range(1, nl):
            sn, H, n = ot.trial_numpy(t, x, 1e9, k, wmax, m, nl, 1.0)
            best = max(best, -sn) if np.isfinite(sn) else best
        np.testing.assert_allclose(out["sn64"][k, 0], best, rtol=1e-7)


def test_float64_gtls_residual_identity():
    """S/N = H/sqrt(res/(n-3)) of the float64 restatement implies the direct
    residual against gen_trapzoid with the exact shifted time (1e-9)."""
    from tests.test_oracle_tls_pins import gen_trapzoid
    rng = np.random.default_rng(22)
    nt = 600
    t = (np.arange(nt) * 0.0204).astype(np.float32)
    x = (1e-3 * rng.standard_normal(nt)).astype(np.float32)
    x[(t > 5.0) & (t < 5.12)] += 3e-3
    tu = t.copy()
    tu[300:310] = -1000.0
    nw, nl, wmax, dw, deltat = 8, 6, 0.3, 0.03, np.float32(0.0204)
    nsc = int(wmax / deltat + 3)
    out = og.gtls64(x[:, None], tu[:, None], nw, nl, nsc, wmax, dw, deltat)
    checked = 0
    for k in range(20, nt - 20, 7):
        sn = float(out["sn64"][k, 0])
        if sn <= 0:
            continue
        W, L, H = (float(out[n][k, 0]) for n in ("tlsw", "tlsl", "tlshmax"))
        tau32, xx = _gtls_scoop(x, tu, k, nsc, deltat)
        rows = np.arange(k - nsc // 2, k - nsc // 2 + nsc)
        rows = rows[(rows >= 0) & (rows < nt)]
        T = (np.float32(-1000.0) + (tu[rows] + np.float32(1000.0))).astype(np.float32)
        tau = T[T >= 0].astype(np.float64) - float(deltat) * k
        W32 = np.float32(W)
        lo, hi = np.float32(-0.5 * W32), np.float32(0.5 * W32)
        mean = xx[(tau32 >= lo) & (tau32 <= hi)].mean()
        win = (tau32 > lo) & (tau32 <= hi)
        r = xx[win] - mean - gen_trapzoid(tau[win], H, W, L, 0.0) * 1.0
        res_direct = float(np.sum(r * r))
        # tlshmax is H rounded to float32: recompute the implied residual from
        # the float64 S/N with that H and compare at the float32 H's precision
        res_implied = (win.sum() - 3) * (H / sn) ** 2
        np.testing.assert_allclose(res_implied, res_direct, rtol=2e-6)
        checked += 1
    assert checked > 30


def test_float32_arithmetic_error_against_float64():
    """The float32 restatement (the reference's arithmetic) against the float64
    one on flipped, near-unit flux: a median of a few 1e-6 in S/N, the same
    best trials almost everywhere."""
    rng = np.random.default_rng(23)
    nt = 1500
    t = (np.arange(nt) * 0.0204).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal(nt)).astype(np.float32)
    x[(t > 9.0) & (t < 9.15)] += 3e-3
    kw = dict(nw=16, nl=10, nsc=27, wmax=0.5, dw=0.03, deltat=np.float32(0.0204))
    a = og.gtls(x[:, None], t[:, None], **kw)
    b = og.gtls64(x[:, None], t[:, None], **kw)
    pos = (a["tlssn"] > 0) & (b["tlssn"] > 0)
    rel = np.abs(a["tlssn"][pos] - b["sn64"][pos]) / b["sn64"][pos]
    assert np.median(rel) < 2e-5 and np.mean(a["tlsw"] == b["tlsw"]) > 0.98
