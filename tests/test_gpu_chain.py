"""The reference's single-transit driver chain on the device against the oracle
chain, stage by stage (kepler/gtls_mockkepler.py:213-376,444): transit flip
2 - lc, median detrend with r = 128 (subtract, gtrap/gfilter.py), gtls over
every epoch, peak statistics of the S/N series (detect_peaks with mpd 48,
ntop 4), and the cut-out around the top peak (picktrap, alpha 1, nx 201).

The device chain stays on the device between stages (the detrended image
feeds gtls, gtls's outputs feed the peak statistics, the detrended image and
the top peak's T0 and W feed the cut-out); r = 128 runs the wide-window
median.  Parity mode is exact at every
stage; the fast gtls is checked against the float64 oracle, and the stages
after it take the oracle's inputs from the device's gtls outputs.
"""
import numpy as np
import pytest

from oracle import gtls as og
from oracle import medfilt as om
from oracle import peakstats as ops
from oracle import picktrap as opt

pytestmark = pytest.mark.gpu


def _image(nt=6000, nq=3, seed=90):
    from gtrap_amd import synth
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * 0.0204
    lc = np.empty((nt, nq), np.float32)
    tu = np.empty((nt, nq), np.float32)
    for q in range(nq):
        x = 1.0 + 1e-3 * rng.standard_normal(nt) + 2e-3 * np.sin(2 * np.pi * t / 23.0 + q)
        x -= synth.trapezoid_dip(t, rng.uniform(20, 100), None, 4e-3, 0.35, 0.04)
        tt = t.copy()
        g0 = int(rng.integers(500, nt - 700))
        tt[g0:g0 + 60] = -1000.0  # a gap
        tt[rng.random(nt) < 0.01] = -1000.0
        lc[:, q] = x
        tu[:, q] = tt
    return (2.0 - lc).astype(np.float32), tu  # transit mode flip


@pytest.mark.parametrize("arith", ["parity", "fast"])
def test_driver_chain_matches_oracle_chain(gpu, arith):
    import torch
    from gtrap_amd import gfilter, gtls, peakstats, picktrap
    lc, tu = _image()
    nt, nq = lc.shape
    # detrend with the driver's window r = 128 (subtract), on the device
    dev_lc = torch.from_numpy(lc).cuda()
    dev_det = torch.empty_like(dev_lc)
    gfilter.medfilt_device(dev_lc, dev_det, nt, nq, 128, "sub")
    det = dev_det.cpu().numpy()
    det_o = om.detrend_scipy(lc, 128, "sub")
    np.testing.assert_array_equal(det.view(np.uint32), det_o.view(np.uint32))
    # gtls on the device image
    prm = gtls.gtls_params(tu[:, 0], 0.1, 0.6, 64, 10)
    outs = [torch.empty(nt * nq, dtype=torch.float32, device="cuda") for _ in range(5)]
    dev_tu = torch.from_numpy(tu).cuda()
    gtls.gtls_device(*outs, dev_det.reshape(-1), dev_tu, nt, nq, 64, 10, prm["nsc"], prm["wmax"],
                     prm["dw"], prm["deltat"], arith=arith)
    torch.cuda.synchronize()
    got = [o.cpu().numpy().reshape(nt, nq) for o in outs]
    args = (64, 10, prm["nsc"], prm["wmax"], prm["dw"], prm["deltat"])
    if arith == "parity":
        want = og.gtls(det_o, tu, *args)
        for g, k in zip(got, ("tlssn", "tlsw", "tlst0", "tlsl", "tlshmax")):
            np.testing.assert_array_equal(g.view(np.uint32), want[k].view(np.uint32), err_msg=k)
    else:
        want = og.gtls64(det_o, tu, *args)
        np.testing.assert_allclose(got[0], want["tlssn"], rtol=3e-7)
        assert np.mean(got[1] == want["tlsw"]) > 0.99
    sn, w, t0, l, h = got
    # peak statistics of every curve's S/N series, on the device outputs
    ps = peakstats.get_peakstats(outs[0], nt, nq, mpd=48, ntop=4, tlst0=outs[2], tlsw=outs[1],
                                 tlsl=outs[3], tlshmax=outs[4])
    for q in range(nq):
        ref = ops.peakstats(sn[:, q], t0=t0[:, q], mpd=48, ntop=4)
        assert ps["npeak"][q] == ref["npeak"] and ps["lent"][q] == ref["lent"]
        np.testing.assert_array_equal(ps["top"][q][:len(ref["top"])], ref["top"])
        for k in ("maxsn", "median", "std", "far", "diff", "pssn"):
            np.testing.assert_allclose(ps[k][q], ref[k], rtol=1e-12, err_msg=k)
        k0 = ref["top"][0]
        assert ps["T0"][q] == t0[k0, q] and ps["W"][q] == w[k0, q]
    # cut-outs around the top peaks from the device image
    lcsx, tx, igapx, prec, status = picktrap.cutouts(dev_det, dev_tu, np.arange(nq), ps["T0"],
                                                     ps["W"], alpha=1, nx=201)
    for q in range(nq):
        rl, _, rg, rp, rs = opt.cutout(det_o[:, q], tu[:, q], float(ps["T0"][q]),
                                       float(ps["W"][q]), alpha=1, nx=201)
        assert bool(prec[q]) == rp and status[q] == rs
        np.testing.assert_allclose(lcsx[q], rl, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(igapx[q], rg, rtol=1e-12, atol=1e-12)
