"""Pins oracle.precision.bls64_cells (the float64 yardstick of the precision
study) against the float64 brute force, a closed form and an invariant."""
import numpy as np

from oracle import bls as ob
from oracle import precision as op


def _exact_grid_case(nq=3, n=2048, seed=0):
    # t*f exact in float32 and float64 (t has 2 fractional bits, f 10), so the
    # float64 fold bins exactly like the reference's float32 fold
    rng = np.random.default_rng(seed)
    t = (np.arange(n) * 0.25).astype(np.float64)
    lc = (1.0 + 1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    tu = np.broadcast_to(t.astype(np.float32)[:, None], (n, nq)).copy()
    tu[rng.random((n, nq)) < 0.05] = -1000.0
    off = lc.astype(np.float64).mean(0).astype(np.float32)
    ntrue = (tu >= 0).sum(0).astype(np.uint32)
    return t, lc, tu, off, ntrue


def test_equals_brute64_on_exact_fold():
    t, lc, tu, off, ntrue = _exact_grid_case()
    nb, kma, kmi, kkmi, nthread = 256, 20, 3, 30, 256
    fmin, df, nf = 0.0625, 2.0 ** -10, 12
    want, _ = ob.geebls_brute64(lc, tu, off, ntrue, nf, nb, kma, kmi, kkmi, fmin, df,
                                nthread=nthread)
    cells = [(q, jf) for q in range(lc.shape[1]) for jf in range(nf)]
    got = op.bls64_cells(lc, t, tu, off, ntrue, nb, kma, kmi, kkmi, fmin, df, nthread, cells)
    np.testing.assert_allclose(got, want.reshape(-1), rtol=1e-12)


def test_noiseless_box_closed_form():
    n, nb = 4096, 512
    t = np.arange(n) * 0.125
    P = 16.0  # t*f and the bin edges are exact binary fractions
    ph = (t / P) % 1.0
    inbox = (ph >= 0.25) & (ph < 0.25 + 8 / nb)
    depth = float(np.float32(1e-3))  # the stored value
    lc = np.where(inbox, -depth, 0.0).astype(np.float32)[:, None]
    tu = t.astype(np.float32)[:, None]
    off = np.zeros(1, np.float32)
    ntrue = np.array([n], np.uint32)
    sr = op.bls64_cells(lc, t, tu, off, ntrue, nb, 16, 2, 10, 1 / P, 1e-4, 256, [(0, 0)])[0]
    kin = inbox.sum()
    np.testing.assert_allclose(sr, depth * np.sqrt(kin / (n - kin)), rtol=1e-12)


def test_sign_flip_invariance():
    t, lc, tu, off, ntrue = _exact_grid_case(seed=4)
    args = (256, 20, 3, 30, 0.07, 3e-4, 256, [(0, 1), (1, 5), (2, 9)])
    a = op.bls64_cells(lc, t, tu, off, ntrue, *args)
    b = op.bls64_cells(-lc, t, tu, -off, ntrue, *args)
    np.testing.assert_allclose(a, b, rtol=1e-12)
