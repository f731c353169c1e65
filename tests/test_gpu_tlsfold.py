"""GPU parity of the periodic trapezoid search against the direct oracle.

The GPU evaluates each trial from float64 prefix sums; the oracle sums every
window directly.  Region membership is decided by the same float64 offset
test, so S/N agrees to float64 rounding (checked at 1e-7 relative) and the
best (W, L, epoch) indices are identical except where the oracle's top two
trials of a period are within 1e-9 relative.
"""
import numpy as np
import pytest

from oracle import tlsfold as ot

pytestmark = pytest.mark.gpu


def _curve(n, seed, mask=0.03, inject=None):
    from gtrap_amd import synth
    rng = np.random.default_rng(seed)
    t = np.arange(n) * 0.0204
    x = 1.0 + 1e-3 * rng.standard_normal(n)
    if inject:
        P, t0, t14, ing, depth = inject
        x -= synth.trapezoid_dip(t, t0, P, depth, t14, ing)
    t = t.astype(np.float32)
    if mask:
        t[rng.random(n) < mask] = -1000.0
    return t, x.astype(np.float32)


def _check(curves, plist, widths, nl, dtep=0.0204):
    from gtrap_amd import tlsfold
    got = tlsfold.tlsfold(curves, plist, widths, nl, dtep, nblocks=64)
    for (t, x), per, g in zip(curves, plist, got):
        want = ot.tlsfold(t, x, per, widths, nl, dtep)
        fin = np.isfinite(want["sn"])
        np.testing.assert_array_equal(np.isfinite(g["sn"]), fin)
        np.testing.assert_allclose(g["sn"][fin], want["sn"][fin], rtol=1e-7)
        same = (g["w"] == want["w"]) & (g["l"] == want["l"]) & (g["e"] == want["e"])
        for p in np.nonzero(~same)[0]:
            # only near-ties may pick a different trial; verify the GPU's pick
            # scores the same as the oracle's best within 1e-9 relative
            sn_g = ot.trial(t, x, per[p], g["e"][p], widths[g["w"][p]], g["l"][p], nl, dtep)[0]
            assert abs(sn_g - want["sn"][p]) <= 1e-9 * abs(want["sn"][p]), (p, g, want)


def test_small_random_curves(gpu):
    widths = np.linspace(0.04, 0.3, 10)
    curves = [_curve(800, 1), _curve(1500, 2, inject=(1.7, 0.4, 0.12, 0.01, 3e-3)),
              _curve(333, 3, mask=0.2)]
    plist = [np.geomspace(0.5, 5.0, 23), np.geomspace(0.6, 9.0, 31), np.geomspace(0.5, 2.2, 7)]
    _check(curves, plist, widths, 10)


def test_chunked_windows_small_periods(gpu):
    """Long curve, short periods: one epoch's window exceeds the LDS tile."""
    widths = np.array([0.1, 0.3])
    curves = [_curve(12000, 4, inject=(0.73, 0.2, 0.1, 0.01, 2e-3))]
    plist = [np.array([0.5, 0.61, 0.73, 0.9])]
    _check(curves, plist, widths, 6)


def test_periods_below_the_widest_window(gpu):
    """Periods below the widest window: one epoch's range holds the folded
    curve more than twice (chunked sweep over images of both neighbouring
    cycles); the widths at or above the period give no trial, those below
    it do."""
    widths = np.array([0.05, 0.1, 0.3])
    curves = [_curve(12000, 41, inject=(0.13, 0.05, 0.04, 0.004, 3e-3))]
    plist = [np.array([0.11, 0.12, 0.13, 0.14])]
    _check(curves, plist, widths, 5)


def test_single_epoch_limit_and_nl_grid(gpu):
    widths = np.array([0.05, 0.2])
    t, x = _curve(400, 5, mask=0.0)
    base = float(t.max())
    _check([(t, x)], [np.array([2.5 * base, 4.0 * base])], widths, 4)


def test_single_curve_grid_matches_oracle(gpu):
    """The single-curve workload (4096 samples, trapezoid at P = 3.7 d) on every 20th
    period of the 1000-period grid plus the true period: the GPU's per-period
    best trials equal the direct oracle's.  (tls.py's S/N is H over the
    per-sample residual RMS, so the periodogram maximum need not sit at the
    injected period; the comparison is with the oracle, not the injection.)"""
    from gtrap_amd import synth, tlsfold
    d = synth.tls_single_curve()
    per = np.sort(np.concatenate([d["periods"][::20], [d["truth"]["P"]]]))
    _check([(d["t"], d["flux"])], [per], d["widths"], 10)
    got = tlsfold.tlsfold([(d["t"], d["flux"])], [per], d["widths"], 10, 0.0204)[0]
    want = ot.tlsfold(d["t"], d["flux"], per, d["widths"], 10, 0.0204)
    assert int(np.nanargmax(got["sn"])) == int(np.nanargmax(want["sn"]))


def test_degenerate_curves_and_periods(gpu):
    """Curves with 3, 2, 1 and no valid samples (no trial: S/N -inf, indices
    -1); periods below one epoch step and below the widest window (W >= P is
    no trial); a period commensurate with the cadence (phases in a few tight
    clusters, large brackets); exact duplicate phases with different fluxes."""
    rng = np.random.default_rng(21)
    widths = np.linspace(0.04, 0.3, 10)
    tiny = []
    for n, masked in ((3, 0), (10, 8), (1, 0), (6, 6)):
        t = (np.arange(n) * 0.0204).astype(np.float32)
        x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
        t[:masked] = -1000.0
        tiny.append((t, x))
    _check(tiny, [np.array([0.5, 0.03, 2.0])] * 4, widths, 10)
    t, x = _curve(3000, 22)
    _check([(t, x)], [np.array([0.015, 0.2, 0.25, 0.31, 0.7])], np.array([0.1, 0.3, 0.6]), 6)
    n = 6000
    tc = (np.arange(n) * 0.0204).astype(np.float32)
    xc = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    _check([(tc, xc)], [np.array([0.0204 * 7, 0.0204 * 25, 0.0204 * 49])], widths, 10)
    td = (np.arange(400) * 0.25).astype(np.float32)  # phases exactly 0, 0.25, ...
    xd = (1.0 + 1e-3 * rng.standard_normal(400)).astype(np.float32)
    _check([(td, xd)], [np.array([0.5, 1.0, 0.75])], np.array([0.1, 0.2, 0.4]), 5)


def test_longest_serving_curve(gpu):
    """A curve of the serving maximum (70000 samples, 2% masked) at a chunked
    short period, a middle one and a long one (few cycles per window)."""
    t, x = _curve(70000, 23, mask=0.02)
    _check([(t, x)], [np.array([0.6, 23.3, 466.0])], np.linspace(0.04, 0.3, 10), 10)


def test_dense_periods_over_many_wide_passes(gpu):
    """Windows wider than the LDS tile at every epoch (70000 samples folded at
    3-8 d): each period takes 4-10 wide passes of up to 42 epochs, each pass
    staged in several chunks, so many (epoch, W, boundary) rows are resolved
    at a chunk's first or last sample; the full 10 W x 10 L grid."""
    t, x = _curve(70000, 24, mask=0.02, inject=(5.0, 1.3, 0.12, 0.01, 2e-3))
    _check([(t, x)], [np.array([3.1, 5.0, 8.2])], np.linspace(0.04, 0.3, 10), 10)


def test_gapped_curve(gpu):
    """Two 13-day sectors with a 1-day mid-sector gap each and 10 days between
    them: at periods near and above the gap lengths whole stretches of phase
    hold no sample, so the phase index has long runs of empty cells and many
    boundaries fall in them or past the staged range's ends."""
    rng = np.random.default_rng(31)
    t = np.arange(0.0, 36.0, 0.0204)
    keep = ((t < 6.0) | ((t >= 7.0) & (t < 13.0)) | ((t >= 23.0) & (t < 29.0)) | (t >= 30.0))
    t = t[keep]
    x = 1.0 + 1e-3 * rng.standard_normal(t.size)
    _check([(t.astype(np.float32), x.astype(np.float32))],
           [np.array([0.9, 1.3, 2.7, 6.6, 11.0, 17.5, 40.0])], np.linspace(0.04, 0.3, 10), 10)


def test_high_snr_expanded_residual(gpu):
    """A trapezoid of the trial grid itself (depth 1e-2) at noise 1e-6, sampled
    at the midpoints of a grid aligned to its region boundaries at every cycle
    (cadence 1/64 d, P = 256 cadences): the model matches up to the noise, so
    the best trial's residual is ~4e-8 of the window's sum of squares and the
    kernel's expanded residual (y2 denA^2 - numB denA U + numB^2 V) cancels by
    seven digits.  S/N (~1e4) must still match the direct oracle within 1e-5,
    with the same best trials."""
    from gtrap_amd import tlsfold
    rng = np.random.default_rng(41)
    dt = 1.0 / 64
    P, W, nl, m, e = 4.0, 2.0, 9, 5, 100
    L = m * (0.45 * W / nl)
    t = (np.arange(25 * 256) + 0.5) * dt
    tn = (t - e * dt + P / 2) % P - P / 2
    g = np.zeros_like(tn)
    g[tn <= -W / 4 - L / 2] = -0.5
    g[tn >= W / 4 + L / 2] = -0.5
    g[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = 0.5
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    g[mB] = (tn[mB] + W / 4) / L
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    g[mD] = -(tn[mD] - W / 4) / L
    x = 1.0 - 1e-2 * g + 1e-6 * rng.standard_normal(t.size)
    t, x = t.astype(np.float32), x.astype(np.float32)
    per = np.array([3.9, 4.0, 4.1])
    widths = np.array([1.0, 1.5, 2.0])
    got = tlsfold.tlsfold([(t, x)], [per], widths, nl, dt, nblocks=16)[0]
    want = ot.tlsfold(t, x, per, widths, nl, dt)
    assert (want["w"][1], want["l"][1], want["e"][1]) == (2, m, e)
    assert want["sn"][1] > 1e3
    np.testing.assert_allclose(got["sn"], want["sn"], rtol=1e-5)
    np.testing.assert_array_equal(got["w"], want["w"])
    np.testing.assert_array_equal(got["l"], want["l"])
    np.testing.assert_array_equal(got["e"], want["e"])
    np.testing.assert_allclose(got["h"], want["h"], rtol=1e-6)


def test_noiseless_aligned_trapezoid_known_answer(gpu):
    """An exact trapezoid (tls.py:2-15) sampled at the midpoints of a grid
    aligned to its region boundaries has zero mean over its window, so the
    mean-removed fit recovers H exactly and the residual is 0: the search must
    pick exactly that trial (S/N infinite or huge) with H = -2^-6, for a
    period above the span of the data (the fold is the identity)."""
    from gtrap_amd import tlsfold
    dt = 1.0 / 64
    W, nl, m, e = 2.0, 9, 5, 400
    L = m * (0.45 * W / nl)
    t0 = e * dt
    k = np.arange(-64, 64)
    t = t0 + (k + 0.5) * dt
    H = -2.0 ** -6
    tn = t - t0
    g = np.zeros_like(tn)
    g[tn <= -W / 4 - L / 2] = -0.5
    g[tn >= W / 4 + L / 2] = -0.5
    g[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = 0.5
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    g[mB] = (tn[mB] + W / 4) / L
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    g[mD] = -(tn[mD] - W / 4) / L
    x = H * g
    t32, x32 = t.astype(np.float32), x.astype(np.float32)
    assert np.array_equal(t32, t) and np.array_equal(x32, x)
    widths = np.array([1.0, 2.0])
    per = np.array([16.0])
    got = tlsfold.tlsfold([(t32, x32)], [per], widths, nl, dt, nblocks=8)[0]
    want = ot.tlsfold(t32, x32, per, widths, nl, dt)
    for r in (got, want):
        assert (r["w"][0], r["l"][0], r["e"][0]) == (1, m, e)
        assert abs(r["h"][0] - H) <= 1e-12 * abs(H)
        assert r["sn"][0] == np.inf or r["sn"][0] > 1e6


def test_nonfinite_flux_skipped(gpu):
    """NaN and +-inf flux values are skipped like masked samples, as in the
    oracle (the windows around them keep their other samples)."""
    t, x = _curve(3000, 51, mask=0.02)
    rng = np.random.default_rng(52)
    bad = rng.choice(3000, 40, replace=False)
    x = x.copy()
    x[bad[:20]] = np.nan
    x[bad[20:30]] = np.inf
    x[bad[30:]] = -np.inf
    x[0] = np.nan  # the first sample (the flux reference) is skipped too
    _check([(t, x)], [np.geomspace(0.5, 8.0, 17)], np.linspace(0.04, 0.3, 10), 10)
    tm = t.copy()
    tm[~np.isfinite(x)] = -1000.0
    from gtrap_amd import tlsfold
    a = tlsfold.tlsfold([(t, x)], [np.geomspace(0.5, 8.0, 17)], np.linspace(0.04, 0.3, 10))[0]
    b = tlsfold.tlsfold([(tm, x)], [np.geomspace(0.5, 8.0, 17)], np.linspace(0.04, 0.3, 10))[0]
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_clustered_phases_larger_than_a_sort_chunk(gpu):
    """20000 samples sharing 2 exact time stamps (so 2 phases at every
    period) plus a regular curve: one phase bucket holds 10000 samples, more
    than an LDS sort chunk, and is ordered by the workgroup's bitonic sort."""
    rng = np.random.default_rng(61)
    tr = np.arange(3000) * 0.0204
    t = np.concatenate([np.full(10000, 12.3), tr, np.full(10000, 40.1)])
    x = 1.0 + 1e-3 * rng.standard_normal(t.size)
    _check([(t.astype(np.float32), x.astype(np.float32))], [np.array([0.7, 1.3, 2.9, 9.7])],
           np.linspace(0.04, 0.3, 10), 10)


def test_curve_longer_than_the_workspace_slot_is_refused(gpu):
    """A curve longer than nmax is not searched (S/N NaN, indices -2); the
    others of the batch are."""
    import torch
    from gtrap_amd import _native, tlsfold
    curves = [_curve(500, 71), _curve(900, 72)]
    per = [np.array([0.7, 1.9]), np.array([0.8, 2.2, 3.3])]
    b = tlsfold.pack(curves, per)
    b.nmax = 600  # the second curve does not fit its slot
    out = tlsfold.tlsfold_device(b, np.linspace(0.04, 0.3, 10), 10, 0.0204, nblocks=4,
                                 workspace=tlsfold.Workspace())
    torch.cuda.synchronize()
    sn = out["sn"].cpu().numpy()
    w = out["w"].cpu().numpy()
    assert np.all(np.isnan(sn[2:])) and np.all(w[2:] == -2)
    want = ot.tlsfold(*curves[0], per[0], np.linspace(0.04, 0.3, 10), 10, 0.0204)
    np.testing.assert_allclose(sn[:2], want["sn"], rtol=1e-7)


def test_detection_statistic_matches_oracle(gpu):
    """stat="dchi2" (least-squares matched filter): S/N-like value, H*, and
    the best (W, L, epoch) against the oracle, tile and chunked paths."""
    from gtrap_amd import tlsfold
    widths = np.linspace(0.04, 0.3, 10)
    curves = [_curve(1500, 81, inject=(1.7, 0.4, 0.12, 0.01, 3e-3)), _curve(12000, 82)]
    plist = [np.geomspace(0.6, 9.0, 21), np.array([0.5, 0.73, 3.3, 60.0])]
    got = tlsfold.tlsfold(curves, plist, widths, 10, 0.0204, nblocks=64, stat="dchi2")
    for (t, x), per, g in zip(curves, plist, got):
        want = ot.tlsfold(t, x, per, widths, 10, 0.0204, stat="dchi2")
        fin = np.isfinite(want["sn"])
        np.testing.assert_allclose(g["sn"][fin], want["sn"][fin], rtol=1e-7)
        np.testing.assert_allclose(g["h"][fin], want["h"][fin], rtol=1e-7)
        same = (g["w"] == want["w"]) & (g["l"] == want["l"]) & (g["e"] == want["e"])
        for p in np.nonzero(~same)[0]:
            s = ot.trial(t, x, per[p], g["e"][p], widths[g["w"][p]], g["l"][p], 10, 0.0204,
                         stat="dchi2")[0]
            assert abs(s - want["sn"][p]) <= 1e-9 * abs(want["sn"][p])


def test_tile_geometry_changes_no_result(gpu):
    """Phase-index cells and window-table runs leave every bit unchanged; the
    tile epochs and the bucket size (which sets the staged sample ranges)
    move S/N by float64 rounding only."""
    from gtrap_amd import autotune, tlsfold
    widths = np.linspace(0.04, 0.3, 10)
    curves = [_curve(9000, 91, inject=(2.2, 0.3, 0.15, 0.015, 3e-3)), _curve(2500, 92)]
    plist = [np.geomspace(0.5, 60.0, 40), np.geomspace(0.5, 16.0, 25)]
    try:
        autotune.set_tiles(144, 2, 2, 0)
        base = tlsfold.tlsfold(curves, plist, widths, 10, 0.0204, nblocks=128)
        for geo in ((144, 1, 2, 0), (144, 2, 2, 24), (144, 1, 2, 36)):
            autotune.set_tiles(*geo)
            got = tlsfold.tlsfold(curves, plist, widths, 10, 0.0204, nblocks=128)
            for g, b in zip(got, base):
                for k in g:
                    np.testing.assert_array_equal(g[k], b[k], err_msg=f"{geo} {k}")
        for geo in ((48, 2, 2, 0), (96, 1, 4, 24), (144, 2, 1, 0), (144, 1, 8, 36)):
            autotune.set_tiles(*geo)
            got = tlsfold.tlsfold(curves, plist, widths, 10, 0.0204, nblocks=128)
            for g, b in zip(got, base):
                # the kernel's own agreement with the direct oracle (1e-7)
                np.testing.assert_allclose(g["sn"], b["sn"], rtol=1e-7)
                assert np.mean(g["w"] == b["w"]) > 0.95
    finally:
        autotune.apply_tiles(force=True)
