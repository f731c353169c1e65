"""Caller-side parameter helpers pinned to the values the reference's drivers
compute (kepler/gtls_mockkepler.py:235-256,295; the notebooks' BLS cell)."""
import numpy as np

from gtrap_amd import geebls, gtls
from gtrap_amd.tlsfold import period_grid


def test_bls_params_notebook_values():
    # qma 0.1, qmi 0.01, nb 1024, n 16384: kma = int(102.4) + 1, kmi = int(10.24),
    # kkmi = int(163.84), shared = 4 (nb + kma) 2 + 4 nb
    assert geebls.bls_params(16384, 1024, 0.01, 0.1) == (103, 10, 163, 13112)
    assert geebls.bls_params(8192, 512, 0.02, 0.05) == (26, 10, 163, 6352)


def test_gtls_params_kepler_setting():
    # 72000 samples at 0.0204 d with a masked head and tail: deltat from the
    # valid range, nw 1024 widths over 1-6 d, nl 20
    tu = (np.arange(72000) * 0.0204).astype(np.float32)
    tu[:5] = -1000.0
    tu[-3:] = -1000.0
    p = gtls.gtls_params(tu, 1.0, 6.0, 1024, 20)
    assert abs(float(p["deltat"]) - 0.0204) < 1e-6
    assert p["nsc"] == 297                       # int(6/0.0204 + 3) = int(297.1)
    np.testing.assert_allclose(float(p["dw"]), 5.0 / 1023, rtol=1e-7)
    assert p["shared"] == (2 * 297 + 1024 + 2) * 4 == 6480


def test_period_grid_ends_and_counts():
    g = period_grid(2048 * 0.0204, n=2048)
    assert len(g) == 500 and g[0] == 0.5
    np.testing.assert_allclose(g[-1], 2048 * 0.0204 / 3, rtol=1e-15)
    g = period_grid(70000 * 0.0204, n=70000)
    assert len(g) == 4000
    np.testing.assert_allclose(g[-1], 476.0, rtol=1e-12)
    # geometric midpoint of the lengths -> geometric mean of the counts
    assert len(period_grid(100.0, n=int(round(np.sqrt(2048 * 70000))))) == 1414
    np.testing.assert_allclose(np.diff(np.log(g)), np.log(g[1] / g[0]), rtol=1e-9)
