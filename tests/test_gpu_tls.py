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


def _best(H, res, n, xs=None, with_sn=True):
    import torch
    from gtrap_amd import _native
    d = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()
    Hd, rd, nd = d(H, np.float64), d(res, np.float64), d(n, np.int32)
    xd = d(xs, np.float64) if xs is not None else None
    sn = torch.full((len(H),), 7.0, dtype=torch.float64, device="cuda")
    best = torch.zeros(4, dtype=torch.float64, device="cuda")
    p = _native.ptr
    _native.call("gtrap_tlsfit_best", p(Hd), p(rd), p(nd), p(xd), len(H),
                 p(sn) if with_sn else None, p(best), _native.stream_handle())
    torch.cuda.synchronize()
    return sn.cpu().numpy(), best.cpu().numpy()


def _numpy_choice(H, res, n):
    with np.errstate(all="ignore"):
        sn = -H / np.sqrt(res / (n.astype(np.int32) - 3))
    try:
        return sn, int(np.nanargmax(sn))
    except ValueError:
        return sn, -1


@pytest.mark.parametrize("case", ["random", "special", "ties", "neginf", "allnan", "one"])
def test_tlsfit_choice_matches_numpy(gpu, case):
    """tls.py:100-110: the device S/N equals numpy's bit for bit (true
    division and a correctly rounded sqrt), and the index is np.nanargmax's
    (NaN as -inf, first maximum); NaN only gives -1 (numpy raises)."""
    rng = np.random.default_rng(5)
    m = 5000
    H = rng.normal(0, 1e-3, m)
    res = rng.uniform(1e-8, 1e-3, m)
    n = rng.integers(0, 400, m).astype(np.int32)
    if case == "special":
        H[:50] = 0.0
        res[40:90] = 0.0
        res[90:120] = -1e-6
        n[120:200] = 3
        n[200:220] = 2
        H[220:230] = np.inf
    elif case == "ties":
        H[:] = -1e-3
        res[:] = 1e-4
        n[:] = 100
        H[[4000, 17, 2500]] = -2e-3  # the same S/N three times: the first index wins
    elif case == "neginf":
        H[:] = np.nan
        H[[3000, 1234]] = 1.0
        res[[3000, 1234]] = 0.0  # -1/0 = -inf: the maximum is -inf, at the first of them
        n[[3000, 1234]] = 10
        # every S/N is NaN or -inf: nanargmax counts NaN as -inf and returns the
        # first index, trial 0 (a NaN trial), and so must the device
    elif case == "allnan":
        H[:] = np.nan
    elif case == "one":
        H, res, n = H[:1], res[:1], n[:1]
    xs = rng.normal(0, 1, len(H))
    sn, best = _best(H, res, n, xs)
    want_sn, want_i = _numpy_choice(H, res, n)
    nan = np.isnan(want_sn)
    np.testing.assert_array_equal(np.isnan(sn), nan)
    np.testing.assert_array_equal(sn[~nan].view(np.uint64), want_sn[~nan].view(np.uint64))
    assert int(best[0]) == want_i, (case, best[0], want_i)
    if want_i >= 0:
        np.testing.assert_array_equal(best[2:], [H[want_i], xs[want_i]])  # NaN == NaN here
    # without the S/N array and without xsum: same choice
    _, b2 = _best(H, res, n, None, with_sn=False)
    assert int(b2[0]) == want_i and (want_i < 0 or b2[3] == 0.0)


def test_tlsfit_repeated_refits_reuse_buffers(gpu):
    """Refits of one curve on changing grids (the reference's refit loop,
    works/sp/gtls_tesslc.py:299-314) through the reused buffers: each equals
    the choice numpy makes on that grid's device outputs."""
    from gtrap_amd import tls
    rng = np.random.default_rng(8)
    t = np.arange(4000) * (2.0 / 1440.0)
    x = 1.0 + 5e-4 * rng.standard_normal(4000) + tls.gen_trapzoid(t, -2e-3, 0.3, 0.05, 3.1)
    x[rng.random(4000) < 0.01] = np.nan
    for c, nt in ((3.1, 10), (1.7, 7), (4.2, 13), (3.1, 10)):
        tl = np.linspace(c - 0.1, c + 0.1, nt)
        Wl = np.linspace(0.297, 0.303, 10)
        Ll = np.linspace(0.0, 0.15, 20)
        Hd, rd, nd, _ = (a.cpu().numpy() for a in tls.grid_device(t, x, Wl, Ll, tl))
        with np.errstate(all="ignore"):
            sn = -Hd / np.sqrt(rd / (nd - 3))
        w, l, k = np.unravel_index(np.nanargmax(sn), sn.shape)
        xpre, H, W, L, T0, off = tls.tlsfit(t, x, Wl, Ll, tl)
        assert (W, L, T0, H) == (Wl[w], Ll[l], tl[k], Hd[w, l, k])
        with np.errstate(all="ignore"):
            np.testing.assert_array_equal(xpre, tls.gen_trapzoid(t, H, W, L, T0) + off)


def test_tlsfit_unsorted_time(gpu):
    """Shuffled samples: the grid scans every sample and the model curve takes
    the reference's mask form; the choice and xpre equal the sorted run's."""
    from gtrap_amd import tls
    rng = np.random.default_rng(9)
    t = np.sort(rng.uniform(0, 20, 3000))
    x = 1.0 + 1e-3 * rng.standard_normal(3000) + tls.gen_trapzoid(t, -4e-3, 1.0, 0.2, 9.1)
    Wl = np.linspace(0.8, 1.2, 5)
    Ll = np.linspace(0.0, 0.4, 5)
    tl = np.linspace(8.5, 9.7, 25)
    a = tls.tlsfit(t, x, Wl, Ll, tl)
    perm = rng.permutation(3000)
    b = tls.tlsfit(t[perm], x[perm], Wl, Ll, tl)
    assert a[2:5] == b[2:5] and abs(a[4] - 9.1) < 0.1
    np.testing.assert_allclose([a[1], a[5]], [b[1], b[5]], rtol=1e-9)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(b[0], tls.gen_trapzoid(t[perm], b[1], b[2], b[3], b[4]) + b[5])
