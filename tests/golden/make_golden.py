"""Writes tests/golden/golden.npz: small seeded inputs and the oracle's outputs
for every hot-path kernel, plus the known-answer case of gtrap/tls.py:113-122.

The expected outputs come from the pinned oracle (oracle/), never from the
reference, which is not run.  Regenerate with
``python tests/golden/make_golden.py`` (after ``make -C oracle``); the test
tests/test_golden.py checks that the oracle still reproduces every array.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import bls as ob  # noqa: E402
from oracle import gtls as og  # noqa: E402
from oracle import medfilt as om  # noqa: E402
from oracle import tlsfold as ot  # noqa: E402

OUT = Path(__file__).resolve().with_name("golden.npz")


def kat_trapezoid(t, H, W, L, t0):
    """tls.py:2-15's model written out for the known-answer case."""
    tn = t - t0
    x = np.zeros(len(tn))
    x[tn <= -W / 4 - L / 2] = -H / 2
    x[tn >= W / 4 + L / 2] = -H / 2
    x[(tn < W / 4 - L / 2) & (tn >= -W / 4 + L / 2)] = H / 2
    mB = (tn > -W / 4 - L / 2) & (tn <= -W / 4 + L / 2)
    x[mB] = H / L * (tn[mB] + W / 4)
    mD = (tn < W / 4 + L / 2) & (tn >= W / 4 - L / 2)
    x[mD] = -H / L * (tn[mD] - W / 4)
    return x


def main():
    rng = np.random.default_rng(20260)
    g = {}
    # median detrend: 1024 x 6 with outliers and a constant run
    img = (1.0 + 1e-3 * rng.standard_normal((1024, 6))).astype(np.float32)
    img[rng.random(img.shape) < 0.01] += 0.05
    img[300:360, 2] = np.float32(1.0)
    g["mf_img"] = img
    for r in (3, 16, 64, 100):
        for mode in ("sub", "div"):
            g[f"mf_{mode}_r{r}"] = om.detrend_scipy(img, r, mode)
    # geebls with gaps (tu = -1000), offsets and ntrue
    n, nq = 1024, 3
    lc = (1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    tu = np.tile((np.arange(n) * 0.0204).astype(np.float32)[:, None], (1, nq))
    tu[rng.random((n, nq)) < 0.05] = -1000.0
    off = (np.where(tu >= 0, lc, 0).astype(np.float64).sum(0) / n).astype(np.float32)
    ntrue = (tu >= 0).sum(0).astype(np.uint32)
    bg = dict(nf=8, nb=128, kma=10, kmi=2, kkmi=8, fmin=0.05, df=0.013, nthread=128)
    b, p1, p2 = ob.geebls(lc, tu, off, ntrue, bg["nf"], bg["nb"], bg["kma"], bg["kmi"],
                          bg["kkmi"], bg["fmin"], bg["df"], nthread=bg["nthread"])
    g.update(bls_lc=lc, bls_tu=tu, bls_off=off, bls_ntrue=ntrue, bls_blsp=b, bls_p1=p1, bls_p2=p2,
             bls_grid=np.array([bg[k] for k in ("nf", "nb", "kma", "kmi", "kkmi", "fmin", "df",
                                                "nthread")], np.float64))
    # gtls normal and absolute (transit-flipped flux 2 - lc)
    nt, nq2 = 400, 2
    lc2 = (2.0 - (1.0 + 1e-3 * rng.standard_normal((nt, nq2)))).astype(np.float32)
    tu2 = np.tile((np.arange(nt) * 0.0204).astype(np.float32)[:, None], (1, nq2))
    tu2[rng.random((nt, nq2)) < 0.03] = -1000.0
    nw, nl, wmax, dw, deltat = 8, 5, 0.3, 0.03, 0.0204
    nsc = int(wmax / deltat + 3)
    g.update(gtls_lc=lc2, gtls_tu=tu2,
             gtls_par=np.array([nw, nl, nsc, wmax, dw, deltat], np.float64))
    for mode in ("normal", "absolute"):
        o = og.gtls(lc2, tu2, nw, nl, nsc, wmax, dw, deltat, mode=mode)
        for k, v in o.items():
            g[f"gtls_{mode}_{k}"] = v
    # periodic trapezoid search
    t = (np.arange(700) * 0.0204).astype(np.float32)
    t[rng.random(700) < 0.04] = -1000.0
    x = (1.0 + 1e-3 * rng.standard_normal(700)).astype(np.float32)
    per = np.geomspace(0.5, 4.0, 9)
    widths = np.array([0.05, 0.1, 0.2, 0.3])
    o = ot.tlsfold(t, x, per, widths, 5, 0.0204)
    g.update(tf_t=t, tf_x=x, tf_per=per, tf_widths=widths, tf_sn=o["sn"], tf_w=o["w"],
             tf_l=o["l"], tf_e=o["e"], tf_h=o["h"])
    # gtrap/tls.py:113-122 known answer: H = -3, W = 30, L = 5, t0 = 50
    tk = np.linspace(0, 100, 132)
    g.update(kat_t=tk, kat_x=kat_trapezoid(tk, -3.0, 30.0, 5.0, 50.0),
             kat_par=np.array([-3.0, 30.0, 5.0, 50.0]))
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(g)} arrays)")


if __name__ == "__main__":
    main()
