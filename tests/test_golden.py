"""The oracle reproduces every committed golden vector (tests/golden/golden.npz,
written by tests/golden/make_golden.py), and the tls.py known-answer case
(gtrap/tls.py:113-122) holds for the fixture's inputs."""
from pathlib import Path

import numpy as np

from oracle import bls as ob
from oracle import gtls as og
from oracle import medfilt as om
from oracle import tlsfold as ot

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")


def test_medfilt_vectors():
    for r in (3, 16, 64, 100):
        for mode in ("sub", "div"):
            np.testing.assert_array_equal(om.detrend_scipy(G["mf_img"], r, mode),
                                          G[f"mf_{mode}_r{r}"])


def test_geebls_vectors():
    nf, nb, kma, kmi, kkmi, fmin, df, nth = G["bls_grid"]
    b, p1, p2 = ob.geebls(G["bls_lc"], G["bls_tu"], G["bls_off"], G["bls_ntrue"], int(nf),
                          int(nb), int(kma), int(kmi), int(kkmi), fmin, df, nthread=int(nth))
    np.testing.assert_array_equal(b, G["bls_blsp"])
    np.testing.assert_array_equal(p1, G["bls_p1"])
    np.testing.assert_array_equal(p2, G["bls_p2"])


def test_gtls_vectors():
    nw, nl, nsc, wmax, dw, deltat = G["gtls_par"]
    for mode in ("normal", "absolute"):
        o = og.gtls(G["gtls_lc"], G["gtls_tu"], int(nw), int(nl), int(nsc), wmax, dw, deltat,
                    mode=mode)
        for k, v in o.items():
            np.testing.assert_array_equal(v, G[f"gtls_{mode}_{k}"], err_msg=f"{mode} {k}")


def test_tlsfold_vectors():
    o = ot.tlsfold(G["tf_t"], G["tf_x"], G["tf_per"], G["tf_widths"], 5, 0.0204)
    for k in ("sn", "w", "l", "e", "h"):
        np.testing.assert_array_equal(o[k], G[f"tf_{k}"])


def test_known_answer():
    H, W, L, t0 = G["kat_par"]
    sn, h, n = ot.trial_numpy(G["kat_t"], G["kat_x"], 1e6, int(t0), W, 10, 27, 1.0, divx=False)
    assert abs(h - H) < 1e-12
