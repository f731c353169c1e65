"""Host-side pieces of gtrap_amd.tls (no GPU): the sorted-time form of the
trapezoid model equals the reference's mask form (tls.py:2-15) bit for bit."""
import numpy as np
import pytest

from gtrap_amd import tls


@pytest.mark.parametrize("seed", range(6))
def test_sorted_trapezoid_equals_mask_form(seed):
    rng = np.random.default_rng(seed)
    W = float(rng.uniform(0.05, 2.0))
    L = 0.0 if seed == 0 else float(rng.uniform(0.0, W / 2.0))
    H = np.float64(rng.uniform(-5e-3, 5e-3))  # a grid value, as tlsfit passes it
    t0 = float(rng.uniform(5.0, 15.0))
    t = np.sort(rng.uniform(0.0, 20.0, 3000))
    # samples exactly on the region boundaries, and repeated times
    edges = t0 + np.array([-W / 4 - L / 2, W / 4 + L / 2, W / 4 - L / 2, -W / 4 + L / 2, 0.0])
    t = np.sort(np.concatenate([t, edges, edges, t[:50]]))
    tn = t - t0
    with np.errstate(all="ignore"):
        want = tls.gen_trapzoid(t, H, W, L, t0)
        got = tls._trapzoid_sorted(tn, H, W, L)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_sortedness_check():
    assert tls._is_sorted(np.array([0.0, 1.0, 1.0, 2.0]))
    assert not tls._is_sorted(np.array([0.0, 2.0, 1.0]))
    assert not tls._is_sorted(np.array([0.0, np.nan, 2.0]))
    assert not tls._is_sorted(np.array([np.nan, 1.0]))
    assert tls._is_sorted(np.array([3.0]))
