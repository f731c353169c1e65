"""GPU kernels against the committed golden vectors (tests/golden/golden.npz):
median detrend, geebls and gtls bit for bit; the periodic trapezoid search to
1e-7 in S/N with the same best trials; the tls.py known answer (H = -3,
gtrap/tls.py:113-122) through the GPU grid."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_medfilt_golden(gpu):
    import torch
    from gtrap_amd import gfilter
    img = G["mf_img"]
    n, nq = img.shape
    x = _dev(img)
    for r in (3, 16, 64, 100):
        for mode in ("sub", "div"):
            y = torch.empty_like(x)
            gfilter.medfilt_device(x, y, n, nq, r, mode)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(y.cpu().numpy(), G[f"mf_{mode}_r{r}"],
                                          err_msg=f"r={r} {mode}")


def test_geebls_golden(gpu):
    import torch
    from gtrap_amd import geebls
    nf, nb, kma, kmi, kkmi, fmin, df, nth = G["bls_grid"]
    nf = int(nf)
    lc = G["bls_lc"]
    n, nq = lc.shape
    b = torch.empty(nf * nq, dtype=torch.float32, device="cuda")
    p1 = torch.empty(nf * nq, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nf * nq, dtype=torch.int32, device="cuda")
    geebls.geebls_device(b, p1, p2, _dev(lc), _dev(G["bls_tu"]), _dev(G["bls_off"]),
                         _dev(G["bls_ntrue"]), n, nq, nf, int(nb), int(kma), int(kmi), int(kkmi),
                         fmin, df, nthread=int(nth))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.cpu().numpy().reshape(nq, nf), G["bls_blsp"])
    np.testing.assert_array_equal(p1.cpu().numpy().view(np.uint32).reshape(nq, nf), G["bls_p1"])
    np.testing.assert_array_equal(p2.cpu().numpy().view(np.uint32).reshape(nq, nf), G["bls_p2"])


def test_gtls_golden(gpu):
    import torch
    from gtrap_amd import gtls
    nw, nl, nsc, wmax, dw, deltat = G["gtls_par"]
    lc = G["gtls_lc"]
    nt, nq = lc.shape
    for mode in ("normal", "absolute"):
        outs = [torch.empty(nt * nq, dtype=torch.float32, device="cuda") for _ in range(5)]
        gtls.gtls_device(*outs, _dev(lc), _dev(G["gtls_tu"]), nt, nq, int(nw), int(nl), int(nsc),
                         wmax, dw, deltat, mode=mode, arith="parity")
        torch.cuda.synchronize()
        for o, k in zip(outs, ("tlssn", "tlsw", "tlst0", "tlsl", "tlshmax")):
            np.testing.assert_array_equal(o.cpu().numpy().reshape(nt, nq),
                                          G[f"gtls_{mode}_{k}"], err_msg=f"{mode} {k}")


def test_tlsfold_golden(gpu):
    from gtrap_amd import tlsfold
    got = tlsfold.tlsfold([(G["tf_t"], G["tf_x"])], [G["tf_per"]], G["tf_widths"], 5, 0.0204,
                          nblocks=8)[0]
    np.testing.assert_allclose(got["sn"], G["tf_sn"], rtol=1e-7)
    for k in ("w", "l", "e"):
        np.testing.assert_array_equal(got[k], G[f"tf_{k}"], err_msg=k)


def test_tls_known_answer_on_gpu(gpu):
    from gtrap_amd import tls
    H, W, L, t0 = G["kat_par"]
    assert abs(tls.get_Hmax(G["kat_t"], G["kat_x"], W, L, t0) - H) < 1e-10
