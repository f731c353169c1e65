"""GPU candidate cut-outs against the oracle (gtrap/picktrap.py:27-139
restated): prec and status exactly, vectors to 1e-12."""
import numpy as np
import pytest

from oracle import picktrap as opk

pytestmark = pytest.mark.gpu


def _image(n, nq, seed):
    rng = np.random.default_rng(seed)
    tu = np.tile((np.arange(n) * 0.0204).astype(np.float32)[:, None], (1, nq))
    lc = (1.0 + 1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    for q in range(nq):
        for _ in range(6):  # gaps of 1..40 slots
            a = int(rng.integers(0, n - 50))
            tu[a:a + int(rng.integers(1, 40)), q] = -1000.0
    tu[:30, 0] = -1000.0  # a leading gap
    return lc, tu


@pytest.mark.parametrize("nx,alpha,contcrit", [(128, 2, 84), (51, 2, 10), (2001, 1.5, 84)])
def test_cutouts_match_oracle(gpu, nx, alpha, contcrit):
    import torch
    from gtrap_amd import picktrap
    n, nq = 6000, 4
    lc, tu = _image(n, nq, seed=nx)
    rng = np.random.default_rng(1)
    cq = np.repeat(np.arange(nq), 6)
    T0 = rng.uniform(-1.0, n * 0.0204 + 1.0, len(cq))
    T0[0] = 0.2  # near the leading gap
    W = rng.uniform(0.05, 0.5, len(cq))
    lcsx, tx, igapx, prec, status = picktrap.cutouts(torch.from_numpy(lc).cuda(),
                                                     torch.from_numpy(tu).cuda(), cq, T0, W,
                                                     alpha=alpha, nx=nx, contcrit=contcrit)
    for c in range(len(cq)):
        wx, wtx, wg, wp, ws = opk.cutout(lc[:, cq[c]], tu[:, cq[c]], T0[c], W[c], alpha=alpha,
                                         nx=nx, contcrit=contcrit)
        assert status[c] == ws and prec[c] == wp, (c, status[c], ws, prec[c], wp)
        np.testing.assert_array_equal(tx, wtx)
        np.testing.assert_allclose(lcsx[c], wx, rtol=1e-12, atol=1e-12, err_msg=str(c))
        np.testing.assert_allclose(igapx[c], wg, rtol=1e-12, atol=1e-13, err_msg=str(c))


def test_one_curve_form(gpu):
    from gtrap_amd import picktrap
    lc, tu = _image(3000, 1, seed=4)
    x, tx, g, prec = picktrap.pick_Wnormalized_cleaned_lc_direct(lc[:, 0], tu[:, 0], 30.0, 0.3)
    wx, _, wg, wp, _ = opk.cutout(lc[:, 0], tu[:, 0], 30.0, 0.3)
    np.testing.assert_allclose(x, wx, rtol=1e-12, atol=1e-12)
    assert prec == wp
    with pytest.raises(IndexError):
        picktrap.pick_Wnormalized_cleaned_lc_direct(lc[:, 0], tu[:, 0], 1e6, 0.3)
