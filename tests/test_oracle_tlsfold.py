"""Pins the periodic trapezoid oracle (the fold extension of gtrap/tls.py)."""
import numpy as np
import pytest

from oracle import tlsfold as ot


def gen_trapzoid(t, H, W, L, t0):
    """Noiseless trapezoid of tls.py:2-15 (total span W/2+L, flat bottom H/2)."""
    tn = t - t0
    x = np.zeros(len(tn))
    x[tn <= -W / 4 - L / 2] = -H / 2
    x[tn >= W / 4 + L / 2] = -H / 2
    x[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = H / 2
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    x[mB] = H / L * (tn[mB] + W / 4)
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    x[mD] = -H / L * (tn[mD] - W / 4)
    return x


def test_known_answer_tls_py():
    """tls.py:113-122: noiseless trapezoid H=-3, W=30, L=5 at t0=50 on
    t=linspace(0,100,132) -> get_Hmax (no mean subtraction) returns H.  A period
    far above the baseline makes the fold the identity; W=30, nl=27, m=10 gives
    L=5.  With the mean subtracted (get_chi2Hmax, divx=True) the estimate is not
    H for a model that is not zero-mean over the window; both paths agree."""
    t = np.linspace(0, 100, 132)
    x = gen_trapzoid(t, -3.0, 30.0, 5.0, 50.0)
    sn, H, n = ot.trial(t.astype(np.float32), x.astype(np.float32), 1e6, 50, 30.0, 10, 27, 1.0,
                        divx=False)
    # float32 storage of t and x limits agreement to ~1e-6
    assert abs(H + 3.0) < 1e-5
    sn2, H2, n2 = ot.trial_numpy(t, x, 1e6, 50, 30.0, 10, 27, 1.0, divx=False)
    assert abs(H2 + 3.0) < 1e-12 and n2 == n
    a = ot.trial(t.astype(np.float32), x.astype(np.float32), 1e6, 50, 30.0, 10, 27, 1.0)
    b = ot.trial_numpy(t.astype(np.float32), x.astype(np.float32), 1e6, 50, 30.0, 10, 27, 1.0)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12)


def _noisy(n=600, seed=0, mask=True):
    rng = np.random.default_rng(seed)
    t = (np.arange(n) * 0.0204).astype(np.float32)
    x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    if mask:
        t[rng.random(n) < 0.05] = -1000.0
    return t, x


def test_direct_c_equals_tls_py_masks():
    t, x = _noisy()
    rng = np.random.default_rng(1)
    for _ in range(200):
        P = rng.uniform(0.5, 8.0)
        e = int(rng.integers(0, int(np.ceil(P / 0.0204))))
        W = rng.uniform(0.04, 0.3)
        m = int(rng.integers(1, 10))
        a = ot.trial(t, x, P, e, W, m, 10, 0.0204)
        b = ot.trial_numpy(t, x, P, e, W, m, 10, 0.0204)
        assert a[2] == b[2]
        np.testing.assert_allclose(a[0], b[0], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(a[1], b[1], rtol=1e-9, atol=1e-15)


def test_l_zero_is_nan_and_skipped():
    t, x = _noisy(mask=False)
    sn, H, n = ot.trial(t, x, 3.0, 10, 0.2, 0, 10, 0.0204)
    assert np.isnan(sn) and np.isnan(H)


def test_offset_and_scale_invariance():
    t, x = _noisy(seed=2)
    args = (2.7, 40, 0.2, 4, 10, 0.0204)
    s0, H0, _ = ot.trial_numpy(t, x, *args)
    s1, H1, _ = ot.trial_numpy(t, (3.0 * x.astype(np.float64) + 5.0), *args)
    np.testing.assert_allclose(s1, s0, rtol=1e-7)
    np.testing.assert_allclose(H1, 3.0 * H0, rtol=1e-7)


def test_reduces_to_single_epoch_when_period_exceeds_baseline():
    t, x = _noisy(seed=3)
    base = float(t.max())
    for e, W, m in [(100, 0.2, 3), (250, 0.1, 7)]:
        a = ot.trial(t, x, 3 * base, e, W, m, 10, 0.0204)
        d = t.astype(np.float64) - e * 0.0204
        ok = t >= 0
        # unfolded: shift t by t0 and evaluate with the tls.py masks (period -> inf)
        b = ot.trial_numpy(t, x, 1e12, e, W, m, 10, 0.0204)
        np.testing.assert_allclose(a[0], b[0], rtol=1e-10)
        assert ok.sum() > 0 and np.all(np.abs(d[ok]) < 3 * base)


def test_recovers_injected_periodic_trapezoid():
    rng = np.random.default_rng(4)
    n = 3000
    t = np.arange(n) * 0.0204
    P, t0, W, L, H = 2.345, 0.7, 0.2, 0.03, -4e-3
    tn = (t - t0 + P / 2) % P - P / 2
    x = 1.0 + gen_trapzoid(tn, H, W, L, 0.0) + H / 2 + 5e-4 * rng.standard_normal(n)
    periods = np.array([1.9, 2.1, 2.345, 2.6, 3.1])
    widths = np.array([0.1, 0.2, 0.3])
    r = ot.tlsfold(t.astype(np.float32), x.astype(np.float32), periods, widths, 10, 0.0204)
    assert int(np.argmax(r["sn"])) == 2
    assert r["w"][2] == 1 and abs(r["e"][2] * 0.0204 - t0) < 0.03
    assert r["h"][2] < 0


def _brute(t, x, periods, widths, nl, dtep):
    """Best trial per period by evaluating every trial with `trial` (each
    sample folded once, no candidate gathering)."""
    out = []
    for P in periods:
        best = (-np.inf, -1, -1, -1)
        for w, W in enumerate(widths):
            for m in range(nl):
                for e in range(int(np.ceil(P / dtep))):
                    sn = ot.trial(t, x, P, e, W, m, nl, dtep)[0]
                    if sn > best[0]:
                        best = (sn, w, m, e)
        out.append(best)
    return out


def test_window_as_long_as_the_period_is_no_trial():
    """W >= P is NaN in the direct trial; the candidate gathering, whose image
    ranges overlap when the widest window reaches the period, still counts
    every sample once: its best trials equal the brute force over `trial`."""
    t, x = _noisy(n=300, seed=5)
    assert np.isnan(ot.trial(t, x, 0.25, 3, 0.3, 4, 6, 0.0204)[0])
    assert np.isfinite(ot.trial(t, x, 0.35, 3, 0.3, 4, 6, 0.0204)[0])
    widths = np.array([0.1, 0.3, 0.6])
    periods = np.array([0.015, 0.2, 0.25, 0.31, 0.7])
    got = ot.tlsfold(t, x, periods, widths, 6, 0.0204)
    for p, (sn, w, m, e) in enumerate(_brute(t, x, periods, widths, 6, 0.0204)):
        assert (got["w"][p], got["l"][p], got["e"][p]) == (w, m, e)
        if w >= 0:
            assert widths[w] < periods[p]
            np.testing.assert_allclose(got["sn"][p], sn, rtol=1e-12)


def test_too_few_samples_give_no_trial():
    for n, masked in ((1, 0), (3, 0), (10, 8), (6, 6)):
        t = (np.arange(n) * 0.0204).astype(np.float32)
        x = np.ones(n, np.float32)
        t[:masked] = -1000.0
        r = ot.tlsfold(t, x, np.array([0.5, 0.03, 2.0]), np.linspace(0.04, 0.3, 10), 10, 0.0204)
        assert np.all(np.isneginf(r["sn"])) and np.all(r["w"] == -1), (n, masked, r)


def test_numpy_period_search_equals_c_oracle():
    """The numpy restatement (tls.py's masks, vectorised over epochs) and the C
    oracle give the same best trial per period."""
    t, x = _noisy(n=1200, seed=9)
    widths = np.linspace(0.04, 0.3, 10)
    per = np.array([0.5, 0.83, 1.7, 4.4, 9.9, 30.0])
    c = ot.tlsfold(t, x, per, widths, 10, 0.0204)
    for j, P in enumerate(per):
        sn, w, m, e, H = ot.period_numpy(t, x, P, widths, 10, 0.0204)
        assert (w, m, e) == (c["w"][j], c["l"][j], c["e"][j])
        np.testing.assert_allclose([sn, H], [c["sn"][j], c["h"][j]], rtol=1e-9)
