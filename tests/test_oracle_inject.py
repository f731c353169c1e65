"""Pins oracle.inject: Philox4x32-10 against the Random123 known-answer
vectors, the injected profile against gtrap/tls.py:2-15's model, and the
gamma restatement's mean and variance."""
import numpy as np

from oracle import inject as oj


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert oj.philox4x32((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                   0x9B00DBD8)
    assert oj.philox4x32((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E,
                                                                   0xA20BC7C6, 0x6D5451FD)
    assert oj.philox4x32((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
                         (0xA4093822, 0x299F31D0)) == (0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                       0x24126EA1)


def _gen_trapzoid(t, H, W, L, t0):
    # gtrap/tls.py:2-15's model, written out independently for the check
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


def test_profile_is_shifted_tls_model():
    t = np.linspace(-1.0, 1.0, 20001)
    for W, L in ((0.4, 0.05), (0.3, 0.1), (0.2, 0.0001)):
        got = oj.trapezoid_profile(t, 0.0, None, 2e-3, W, L)
        want = _gen_trapzoid(t, 2e-3, W, L, 0.0) + 1e-3
        np.testing.assert_allclose(got, want, atol=1e-15)
    # periodic: the same dip around every transit centre
    P = 1.7
    got = oj.trapezoid_profile(t + 3 * P, 0.0, P, 1e-3, 0.3, 0.05)
    np.testing.assert_allclose(got, oj.trapezoid_profile(t, 0.0, None, 1e-3, 0.3, 0.05),
                               atol=1e-15)


def test_gamma_restatement_moments():
    for a, b in ((2.5, 1.0), (0.6, 3.0)):
        x = oj.gamma(4000, a, b, seed=12345).astype(np.float64)
        assert np.all(np.isfinite(x)) and np.all(x > 0)
        assert abs(x.mean() - a / b) < 4 * np.sqrt(a) / b / np.sqrt(len(x))
        assert abs(x.var() - a / b ** 2) < 0.1 * a / b ** 2
