"""Pins the getstat oracle (gtrap/getstat.py:14-123)."""
import numpy as np

from oracle import getstat as og


def test_closed_form_and_patchwise():
    rng = np.random.default_rng(0)
    nf, nq, df, fmin = 4096 + 100, 3, 1e-3, 0.05
    blsp = rng.random((nq, nf)).astype(np.float32)
    blsp[1, 2000] = 5.0
    blsp[2, 4150] = 9.0  # in the unreduced tail: ignored (getstat.py:73)
    pest, sde, _, _, idx = og.blsstat(blsp, None, None, nf, nq, df, fmin)
    p2, s2 = og.blsstat_patchwise(blsp, nf, nq, df, fmin)
    np.testing.assert_allclose(pest, p2, rtol=1e-15)
    np.testing.assert_allclose(sde, s2, rtol=1e-12)
    assert idx[1] == 2000 and idx[2] < 4096
    np.testing.assert_allclose(pest[1], 1.0 / (2000 * df + fmin), rtol=1e-15)
    # closed form for curve 1
    b = blsp[1, :4096].astype(np.float64)
    ave = b.sum() / nf
    std = np.sqrt((b * b).sum() / nf - ave * ave)
    np.testing.assert_allclose(sde[1], (5.0 - ave) / std, rtol=1e-12)


def test_ties_take_lowest_index():
    nf, nq = 2048, 1
    b = np.zeros((nq, nf), np.float32)
    b[0, 700] = b[0, 1500] = 2.0
    p1 = np.arange(nf, dtype=np.uint32)[None]
    _, _, ph1, _, idx = og.blsstat(b, p1, p1, nf, nq, 1e-3, 0.1)
    assert idx[0] == 700 and ph1[0] == 700
