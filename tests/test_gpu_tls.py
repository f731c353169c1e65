"""GPU tlsfit / get_chi2Hmax (gtrap/tls.py) against the numpy restatement."""
import numpy as np
import pytest

from oracle import tlsfold as ot

pytestmark = pytest.mark.gpu


def test_known_answer(gpu):
    """tls.py:113-122: H = -3 recovered from a noiseless trapezoid (get_Hmax)."""
    from gtrap_amd import tls
    t = np.linspace(0, 100, 132)
    x = tls.gen_trapzoid(t, -3.0, 30.0, 5.0, 50.0)
    assert abs(tls.get_Hmax(t, x, 30.0, 5.0, 50.0) + 3.0) < 1e-12
    H, res, n, xs = tls.get_chi2Hmax(t, x, 30.0, 5.0, 50.0)
    Ho, ro, no, xo = ot.chi2hmax_numpy(t - 50.0, x, 30.0, 5.0)
    np.testing.assert_allclose([H, res, xs], [Ho, ro, xo], rtol=1e-12)
    assert n == no


def test_every_trial_and_tlsfit(gpu):
    from gtrap_amd import tls
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0, 20, 700))
    x = 1.0 + 1e-3 * rng.standard_normal(700) + tls.gen_trapzoid(t, -4e-3, 1.2, 0.2, 8.3) + 2e-3
    x[rng.random(700) < 0.02] = np.nan
    Wl = np.linspace(0.6, 1.8, 5)
    Ll = np.array([0.1, 0.2, 0.3])
    tl = np.linspace(2, 18, 161)
    Hd, rd, nd, xd = (a.cpu().numpy() for a in tls.grid_device(t, x, Wl, Ll, tl))
    for w in range(len(Wl)):
        for l in range(len(Ll)):
            for k in range(0, len(tl), 7):
                Ho, ro, no, xo = ot.chi2hmax_numpy(t - tl[k], x, Wl[w], Ll[l])
                assert nd[w, l, k] == no
                # float64 sums in another association (lanes, then a tree)
                # than numpy's pairwise nansum: 1e-9 covers the ill-conditioned
                # trials (H far outside the data) at 700 samples
                np.testing.assert_allclose([Hd[w, l, k], rd[w, l, k], xd[w, l, k]],
                                           [Ho, ro, xo], rtol=1e-9, atol=1e-14)
    xpre, H, W, L, T0, off = tls.tlsfit(t, x, Wl, Ll, tl)
    with np.errstate(all="ignore"):
        sn = -Hd / np.sqrt(rd / (nd - 3))
    w, l, k = np.unravel_index(np.nanargmax(sn), sn.shape)
    assert (W, L, T0) == (Wl[w], Ll[l], tl[k]) and abs(T0 - 8.3) < 0.15 and H < 0
    np.testing.assert_allclose(xpre, tls.gen_trapzoid(t, H, W, L, T0) + off)


def test_window_search_and_unsorted_time(gpu):
    """Sorted time visits only each trial's window (binary searches with the
    region tests themselves); shuffled samples take the full scan.  Both give
    the numpy restatement's values, and the same counts."""
    from gtrap_amd import tls
    rng = np.random.default_rng(7)
    t = np.sort(rng.uniform(0, 27.4, 5000))
    x = 1.0 + 1e-3 * rng.standard_normal(5000) + tls.gen_trapzoid(t, -3e-3, 0.5, 0.1, 13.0)
    Wl = np.linspace(0.3, 0.7, 4)
    Ll = np.linspace(0.0, 0.3, 5)
    tl = np.concatenate([np.linspace(-0.5, 27.9, 37), [t[0], t[-1], t[100]]])
    perm = rng.permutation(5000)
    a = [v.cpu().numpy() for v in tls.grid_device(t, x, Wl, Ll, tl)]
    b = [v.cpu().numpy() for v in tls.grid_device(t[perm], x[perm], Wl, Ll, tl)]
    np.testing.assert_array_equal(a[2], b[2])
    for u, v in zip((a[0], a[1], a[3]), (b[0], b[1], b[3])):
        np.testing.assert_allclose(u, v, rtol=1e-9, atol=1e-14)
    for w in range(len(Wl)):
        for l in range(1, len(Ll)):
            for k in range(0, len(tl), 5):
                Ho, ro, no, xo = ot.chi2hmax_numpy(t - tl[k], x, Wl[w], Ll[l])
                assert a[2][w, l, k] == no
                np.testing.assert_allclose([a[0][w, l, k], a[1][w, l, k], a[3][w, l, k]],
                                           [Ho, ro, xo], rtol=1e-9, atol=1e-14)
