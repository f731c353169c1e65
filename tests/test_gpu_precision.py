"""GPU arithmetic variants of geebls (precision study).

The build's own arithmetic (float32 storage, exact bin sums, float32 fold)
must equal gtrap_geebls bit for bit; the float64 variant must equal the
float64 oracle to rounding; every other variant is an error measurement and
is only held to loose bounds.
"""
import numpy as np
import pytest

from oracle import precision as op

pytestmark = pytest.mark.gpu


def _data(seed=11):
    from gtrap_amd import synth
    return synth.precision_batch(nq=6, n=32768, nf=24, seed=seed)


def test_default_variant_equals_geebls(gpu):
    import torch
    from gtrap_amd import geebls, precision
    d = _data()
    g = d["grid"]
    inp = precision.Inputs(d)
    sr, p1, p2 = precision.run_variant(inp, "f32", "exact32", "f32")
    n, nq, nf = inp.n, inp.nq, g["nf"]
    blsp = torch.empty(nq * nf, dtype=torch.float32, device="cuda")
    q1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    q2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_device(blsp, q1, q2, inp.lc, inp.tu, inp.offset, inp.ntrue, n, nq, nf,
                         g["nb"], g["kma"], g["kmi"], g["kkmi"], g["fmin"], g["df"],
                         nthread=g["nthread"])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sr.cpu().numpy(),
                                  blsp.cpu().numpy().reshape(nq, nf).astype(np.float64))
    np.testing.assert_array_equal(p1.cpu().numpy(), q1.cpu().numpy().reshape(nq, nf))
    np.testing.assert_array_equal(p2.cpu().numpy(), q2.cpu().numpy().reshape(nq, nf))


def test_float64_variant_matches_oracle(gpu):
    from gtrap_amd import precision
    d = _data(seed=12)
    g = d["grid"]
    inp = precision.Inputs(d)
    sr, _, _ = precision.run_variant(inp, *precision.TRUTH)
    sr = sr.cpu().numpy()
    rng = np.random.default_rng(3)
    cells = [(int(rng.integers(0, inp.nq)), int(rng.integers(0, g["nf"]))) for _ in range(12)]
    tu = np.broadcast_to(d["t64"].astype(np.float32)[:, None], d["lc"].shape)
    want = op.bls64_cells(d["lc"], d["t64"], tu, d["offset"], d["ntrue"], g["nb"], g["kma"],
                          g["kmi"], g["kkmi"], g["fmin"], g["df"], g["nthread"], cells)
    got = np.array([sr[q, jf] for q, jf in cells])
    np.testing.assert_allclose(got, want, rtol=1e-9)


def test_study_error_bounds(gpu):
    from gtrap_amd import precision
    res = precision.study(_data(seed=13))
    for name, r in res.items():
        if "yardstick" in name:
            continue
        assert r["sr_rel_median"] < 1e-3, (name, r)
    # exact float32 bin sums with a float64 fold: only the float32 rounding
    # of bins and running sums remains
    assert res["f32/exact32/f64"]["sr_rel_max"] < 1e-5
    assert res["f32/twosum32/f64"]["sr_rel_max"] < 1e-5


def test_geebls_f64_mode_is_the_float64_variant(gpu):
    import torch
    from gtrap_amd import geebls, precision
    d = _data(seed=14)
    g = d["grid"]
    inp = precision.Inputs(d)
    want, w1, w2 = precision.run_variant(inp, *precision.TRUTH)
    nq, nf = inp.nq, g["nf"]
    sr = torch.empty(nq * nf, dtype=torch.float64, device="cuda")
    p1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_f64_device(sr, p1, p2, inp.lc, inp.tu, inp.t64, inp.offset, inp.ntrue, inp.n,
                             nq, nf, g["nb"], g["kma"], g["kmi"], g["kkmi"], g["fmin"], g["df"],
                             nthread=g["nthread"])
    np.testing.assert_array_equal(sr.cpu().numpy().reshape(nq, nf), want.cpu().numpy())
    np.testing.assert_array_equal(p1.cpu().numpy().reshape(nq, nf), w1.cpu().numpy())


def test_long_baseline_full_length_curves(gpu):
    """The precision workload at its own scale on a stated subset: 32 curves of
    262144 samples at 2-minute cadence (364 d, t*f up to ~50 cycles), 256
    frequencies around the transit period.  The build's default arithmetic
    equals gtrap_geebls bit for bit, the float64 variant equals the float64
    oracle on sampled cells, and float64-fold variants stay within 1e-5 of
    it while the float32 fold (the reference's) shows its larger error."""
    import torch
    from gtrap_amd import geebls, precision, synth
    d = synth.precision_batch(nq=32, n=262144, nf=256, seed=15)
    g = d["grid"]
    inp = precision.Inputs(d)
    n, nq, nf = inp.n, inp.nq, g["nf"]
    assert n == 262144 and d["t64"][-1] > 363.0
    sr, p1, p2 = precision.run_variant(inp, "f32", "exact32", "f32")
    blsp = torch.empty(nq * nf, dtype=torch.float32, device="cuda")
    q1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    q2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_device(blsp, q1, q2, inp.lc, inp.tu, inp.offset, inp.ntrue, n, nq, nf,
                         g["nb"], g["kma"], g["kmi"], g["kkmi"], g["fmin"], g["df"],
                         nthread=g["nthread"])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sr.cpu().numpy(),
                                  blsp.cpu().numpy().reshape(nq, nf).astype(np.float64))
    np.testing.assert_array_equal(p1.cpu().numpy(), q1.cpu().numpy().reshape(nq, nf))
    truth, _, _ = precision.run_variant(inp, *precision.TRUTH)
    truth = truth.cpu().numpy()
    rng = np.random.default_rng(5)
    cells = [(int(rng.integers(0, nq)), int(rng.integers(0, nf))) for _ in range(10)]
    tu = np.broadcast_to(d["t64"].astype(np.float32)[:, None], d["lc"].shape)
    want = op.bls64_cells(d["lc"], d["t64"], tu, d["offset"], d["ntrue"], g["nb"], g["kma"],
                          g["kmi"], g["kkmi"], g["fmin"], g["df"], g["nthread"], cells)
    np.testing.assert_allclose([truth[q, jf] for q, jf in cells], want, rtol=1e-9)
    res = precision.study(d)
    print({k: (v["sr_rel_median"], v["sr_rel_max"]) for k, v in res.items()
           if "yardstick" not in k}, flush=True)
    for name in ("f32/exact32/f64", "f32/twosum32/f64", "f32/atomic32/f64"):
        assert res[name]["sr_rel_max"] < 1e-5, (name, res[name])
    assert res["f32/f64/f32"]["sr_rel_max"] > 10 * res["f32/exact32/f64"]["sr_rel_max"]


def _tls_precision_workload(ncurve=8, nper=16, seed=16):
    """Full-length curves of the long-baseline workload (262144 samples at
    2 minutes, 364 d, trapezoid depth 1e-4 at P = 7.3 d, SNR 7-15) and a
    period grid around P that resolves every transit."""
    from gtrap_amd import synth
    d = synth.precision_batch(nq=ncurve, n=262144, nf=8, seed=seed)
    t = d["t64"].astype(np.float32)
    curves = [(t, d["lc"][:, q]) for q in range(ncurve)]
    per = 7.3 + (np.arange(nper) - nper // 2) * 5e-4
    return curves, [per] * ncurve


def test_tls_precision_variants(gpu):
    """The periodic trapezoid search under single-precision arithmetic: the
    production variant is the float64 kernel itself; float64-fold variants
    with binary16 storage or float32 (compensated) sums stay close to it, and
    every variant is measured."""
    import torch
    from gtrap_amd import precision, tlsfold
    curves, plist = _tls_precision_workload()
    widths = np.linspace(0.04, 0.3, 10)
    dt = 2.0 / 1440.0
    inp = precision.TlsInputs(curves, plist)
    prod = precision.tls_run_variant(inp, 0, widths, 10, dt)
    ref = tlsfold.tlsfold(curves, plist, widths, 10, dt)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(prod["sn"].cpu().numpy(), np.concatenate([r["sn"] for r in ref]))
    res = precision.tls_study(curves, plist, widths, 10, dt)
    print({k: v for k, v in res.items()}, flush=True)
    assert res["f32/twosum32/f64"]["sn_rel_max"] < 1e-4
    assert res["f16/twosum32/f64"]["sn_rel_median"] < 1e-2
    for name, r in res.items():
        if "yardstick" not in name:
            assert np.isfinite(r["sn_rel_median"]), name
