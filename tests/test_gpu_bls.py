"""GPU parity of geebls / geebls_simple / getstat against the oracle.

The GPU kernel and the C restatement perform the same float32 operations in
the same order on the same exact fixed-point bin sums, so SR and phases must
match bit for bit.
"""
import numpy as np
import pytest

from oracle import bls as ob
from oracle import getstat as og

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _small(n=2048, nq=5, seed=0):
    from gtrap_amd import synth
    d = synth.bls_batch(nq=nq, n=n, seed=seed)
    lc, tu, ntrue = d["lc"], d["tu"], d["ntrue"]
    off = np.where(tu >= 0, lc, 0).astype(np.float64).sum(0) / n
    return lc, tu, off.astype(np.float32), ntrue


def _run_geebls(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread):
    import torch
    from gtrap_amd import geebls
    n, nq = lc.shape
    blsp = torch.full((nq * nf,), -1.0, dtype=torch.float32, device="cuda")
    p1 = torch.zeros(nq * nf, dtype=torch.int32, device="cuda")
    p2 = torch.zeros(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_device(blsp, p1, p2, _dev(lc), _dev(tu), _dev(off), _dev(ntrue), n, nq, nf, nb,
                         kma, kmi, kkmi, fmin, df, nthread=nthread)
    torch.cuda.synchronize()
    return (blsp.cpu().numpy().reshape(nq, nf), p1.cpu().numpy().view(np.uint32).reshape(nq, nf),
            p2.cpu().numpy().view(np.uint32).reshape(nq, nf))


@pytest.mark.parametrize("nb,kma,kmi,nthread", [(1024, 22, 10, 512), (256, 30, 2, 128),
                                                (64, 8, 0, 64), (2048, 103, 10, 1024)])
def test_geebls_bit_exact(gpu, nb, kma, kmi, nthread):
    lc, tu, off, ntrue = _small(n=3000, nq=4, seed=nb)
    nf, fmin, df = 37, 1.0 / 30, 0.011
    kkmi = int(3000 * kmi / nb)
    got = _run_geebls(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread)
    want = ob.geebls(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread=nthread)
    for g, w, name in zip(got, want, ("blsp", "phase1", "phase2")):
        np.testing.assert_array_equal(g, w, err_msg=name)


def test_geebls_simple_bit_exact(gpu):
    import torch
    from gtrap_amd import geebls_simple
    rng = np.random.default_rng(1)
    n, nq, nf, nb, kma, kmi, kkmi = 4096, 3, 50, 512, 40, 3, 20
    t = (np.arange(n) * 0.0204).astype(np.float32)
    lc = (1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    blsp = torch.zeros(nq * nf, dtype=torch.float32, device="cuda")
    geebls_simple.geebls_simple_device(blsp, _dev(lc), _dev(t), n, nq, nf, nb, kma, kmi, kkmi,
                                       0.05, 0.013, nthread=256)
    torch.cuda.synchronize()
    want, _, _ = ob.geebls(lc, t, None, None, nf, nb, kma, kmi, kkmi, 0.05, 0.013, nthread=256,
                           simple=True)
    np.testing.assert_array_equal(blsp.cpu().numpy().reshape(nq, nf), want)


def test_geebls_dropin_module_and_getstat(gpu):
    """Reference call shape (include/geebls.h:1 via gbls_module) then get_blsstat."""
    from gtrap_amd import driver as cuda
    from gtrap_amd import geebls, getstat
    lc, tu, off, ntrue = _small(n=4096, nq=3, seed=7)
    n, nq = lc.shape
    nf, nb, qmi, qma = 2048, 1024, 0.01, 0.1
    kma, kmi, kkmi, shared = geebls.bls_params(n, nb, qmi, qma)
    fmin, df = 1.0 / 30, 3e-4
    dev = {k: cuda.mem_alloc(v.nbytes) for k, v in
           dict(lc=lc, tu=tu, off=off, ntrue=ntrue).items()}
    for k, v in dict(lc=lc, tu=tu, off=off, ntrue=ntrue).items():
        cuda.memcpy_htod(dev[k], v)
    blsp = cuda.mem_alloc(4 * nf * nq)
    p1 = cuda.mem_alloc(4 * nf * nq)
    p2 = cuda.mem_alloc(4 * nf * nq)
    k = geebls.gbls_module().get_function("geebls")
    k(blsp, p1, p2, dev["lc"], dev["tu"], dev["off"], dev["ntrue"], np.uint32(n), np.uint32(nf),
      np.uint32(nb), np.uint32(kma), np.uint32(kmi), np.uint32(kkmi), np.float32(fmin),
      np.float32(df), block=(512, 1, 1), grid=(nf, nq), shared=shared)
    hb = np.zeros(nf * nq, np.float32)
    h1 = np.zeros(nf * nq, np.uint32)
    h2 = np.zeros(nf * nq, np.uint32)
    cuda.memcpy_dtoh(hb, blsp)
    cuda.memcpy_dtoh(h1, p1)
    cuda.memcpy_dtoh(h2, p2)
    fs = (0, 96)
    want, w1, w2 = ob.geebls(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, fmin, df, nthread=512,
                             fs=fs)
    np.testing.assert_array_equal(hb.reshape(nq, nf)[:, :96], want[:, :96])
    np.testing.assert_array_equal(h1.reshape(nq, nf)[:, :96], w1[:, :96])
    np.testing.assert_array_equal(h2.reshape(nq, nf)[:, :96], w2[:, :96])
    pest, sde, pb1, pb2 = getstat.get_blsstat(blsp, p1, p2, nf, nq, df, fmin)
    opest, osde, op1, op2, _ = og.blsstat(hb, h1, h2, nf, nq, df, fmin)
    np.testing.assert_array_equal(pest, opest)
    np.testing.assert_allclose(sde, osde, rtol=1e-12)
    np.testing.assert_array_equal(pb1, op1)
    np.testing.assert_array_equal(pb2, op2)


def test_bls_batch_full_size_subset(gpu):
    """1024 curves x 16384 samples, 2048 frequencies, nb 1024, kma 22: the whole
    batch on the GPU, a subset of (curve, frequency) cells against the oracle,
    and box recovery on the injected curves."""
    import torch
    from gtrap_amd import geebls, getstat, synth
    d = synth.bls_batch()
    g = synth.bls_batch_grid()
    lc, tu, ntrue = d["lc"], d["tu"], d["ntrue"]
    n, nq = lc.shape
    off = (np.where(tu >= 0, lc, 0).astype(np.float64).sum(0) / n).astype(np.float32)
    nf = g["nf"]
    blsp = torch.empty(nq * nf, dtype=torch.float32, device="cuda")
    p1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_device(blsp, p1, p2, _dev(lc), _dev(tu), _dev(off), _dev(ntrue), n, nq, nf,
                         g["nb"], g["kma"], g["kmi"], g["kkmi"], g["fmin"], g["df"],
                         nthread=g["nthread"])
    pest, sde, _, _, _ = getstat.blsstats_device(blsp, p1, p2, nf, nq, g["df"], g["fmin"])
    torch.cuda.synchronize()
    hb = blsp.cpu().numpy().reshape(nq, nf)
    rng = np.random.default_rng(5)
    for q in rng.choice(nq, 6, replace=False):
        fs = sorted(rng.choice(nf - 8, 1))[0]
        want, _, _ = ob.geebls(lc, tu, off, ntrue, nf, g["nb"], g["kma"], g["kmi"], g["kkmi"],
                               g["fmin"], g["df"], nthread=g["nthread"], qs=(q, q + 1),
                               fs=(fs, fs + 8))
        np.testing.assert_array_equal(hb[q, fs:fs + 8], want[q, fs:fs + 8])
    opest, osde, _, _, _ = og.blsstat(hb, None, None, nf, nq, g["df"], g["fmin"])
    np.testing.assert_array_equal(pest.cpu().numpy(), opest)
    np.testing.assert_allclose(sde.cpu().numpy(), osde, rtol=1e-10)
    assert np.all(np.isfinite(sde.cpu().numpy()))


def test_box_recovery_on_resolved_grid(gpu):
    """A box transit on a frequency grid fine enough for the baseline
    (df * T << duty cycle) is recovered at the injected period."""
    import torch
    from gtrap_amd import geebls, getstat, synth
    rng = np.random.default_rng(21)
    n, nq = 8192, 4
    t = np.arange(n) * 0.0204
    T = t[-1]
    lc = np.empty((n, nq), np.float32)
    Ps = rng.uniform(2.0, 8.0, nq)
    for q in range(nq):
        f = 1e-3 * rng.standard_normal(n)
        f -= synth.box_dip(t, rng.uniform(0, Ps[q]), Ps[q], 3e-3, 0.03 * Ps[q])
        lc[:, q] = f
    tu = np.tile(t.astype(np.float32)[:, None], (1, nq))
    ntrue = np.full(nq, n, np.uint32)
    off = lc.astype(np.float64).mean(0).astype(np.float32)
    fmin, fmax = 1 / 10.0, 1 / 1.5
    df = 0.2 * 0.03 / T
    nf = int((fmax - fmin) / df) // 1024 * 1024
    nb, kma, kmi = 1024, 60, 10
    blsp = torch.empty(nq * nf, dtype=torch.float32, device="cuda")
    p1 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nq * nf, dtype=torch.int32, device="cuda")
    geebls.geebls_device(blsp, p1, p2, _dev(lc), _dev(tu), _dev(off), _dev(ntrue), n, nq, nf, nb,
                         kma, kmi, int(n * kmi / nb), fmin, df, nthread=512)
    pest, sde, _, _, _ = getstat.blsstats_device(blsp, p1, p2, nf, nq, df, fmin)
    got = pest.cpu().numpy()
    for q in range(nq):
        assert min(abs(got[q] / Ps[q] * h - 1) for h in (1, 2, 0.5)) < 2e-3, (got[q], Ps[q])
    assert np.all(sde.cpu().numpy() > 5)


def test_invalid_arguments_raise(gpu):
    lc, tu, off, ntrue = _small(n=512, nq=1)
    with pytest.raises(ValueError):
        _run_geebls(lc, tu, off, ntrue, 4, 1000, 8, 1, 5, 0.1, 0.01, 500)  # nb not pow2
    with pytest.raises(ValueError):
        _run_geebls(lc, tu, off, ntrue, 4, 1024, 600, 1, 5, 0.1, 0.01, 512)  # kma > nthread


def test_workspace_path_matches_in_place(gpu):
    """The curve-major path (gtrap_geebls_ws / _simple_ws: image copied through
    the workspace, one scale per curve) returns the bits of the in-place entry
    points, on ragged sizes (two 64-curve transpose tiles, a partial sample
    tile), with NaN fluxes and masked times in the batch."""
    import torch
    from gtrap_amd import geebls, geebls_simple
    lc, tu, off, ntrue = _small(n=3001, nq=67, seed=11)
    lc = lc.copy()
    lc[17, 3] = np.nan
    lc[:, 5] = np.nan  # a curve of NaN: scale from 0, as in place
    nf, nb, kma, kmi, nthread = 41, 512, 20, 4, 256
    kkmi = int(3001 * kmi / nb)
    outs = []
    for ws in (False, None, geebls.Workspace()):
        blsp = torch.full((67 * nf,), -1.0, dtype=torch.float32, device="cuda")
        p1 = torch.zeros(67 * nf, dtype=torch.int32, device="cuda")
        p2 = torch.zeros(67 * nf, dtype=torch.int32, device="cuda")
        geebls.geebls_device(blsp, p1, p2, _dev(lc), _dev(tu), _dev(off), _dev(ntrue), 3001, 67,
                             nf, nb, kma, kmi, kkmi, 1.0 / 30, 0.011, nthread=nthread,
                             workspace=ws)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (blsp, p1, p2)])
    for o in outs[1:]:
        for a, b, name in zip(outs[0], o, ("blsp", "phase1", "phase2")):
            np.testing.assert_array_equal(a, b, err_msg=name)
    # against the oracle on the curves without NaN flux (NaN sums have no
    # defined value in the reference's float atomics)
    want = ob.geebls(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, 1.0 / 30, 0.011, nthread=nthread)
    fin = np.isfinite(lc).all(0)
    np.testing.assert_array_equal(outs[1][0].reshape(67, nf)[fin], want[0][fin])
    # simple: common 1-D time, pre-centred flux
    t = (np.arange(3001) * 0.0204).astype(np.float32)
    lcs = (lc - off[None, :]).astype(np.float32)
    res = []
    for ws in (False, None):
        blsp = torch.zeros(67 * nf, dtype=torch.float32, device="cuda")
        geebls_simple.geebls_simple_device(blsp, _dev(lcs), _dev(t), 3001, 67, nf, nb, kma, kmi,
                                           kkmi, 0.05, 0.013, nthread=nthread, workspace=ws)
        torch.cuda.synchronize()
        res.append(blsp.cpu().numpy())
    np.testing.assert_array_equal(res[0], res[1])


def test_workspace_too_small_is_an_error(gpu):
    import torch
    from gtrap_amd import _native
    lc, tu, off, ntrue = _small(n=512, nq=2)
    need = int(_native.load().gtrap_geebls_ws_bytes(512, 2, 0))
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    blsp = torch.zeros(2 * 4, dtype=torch.float32, device="cuda")
    p = _native.ptr
    dl, dt, do, dn = _dev(lc), _dev(tu), _dev(off), _dev(ntrue)  # held across the launches
    args = (p(blsp), None, None, p(dl), p(dt), p(do), p(dn), 512, 2, 4, 256, 8, 1, 5, 0.1, 0.01,
            256)
    with pytest.raises(_native.GtrapError):
        _native.call("gtrap_geebls_ws", *args, p(ws), need - 1, None)
    _native.call("gtrap_geebls_ws", *args, p(ws), need, None)
    torch.cuda.synchronize()
