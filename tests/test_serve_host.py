"""Host-side logic of the serving loop (no GPU): arena allocator and grids."""
import numpy as np

from gtrap_amd.serve import FirstFit, Request
from gtrap_amd.tlsfold import period_grid


def test_first_fit_alloc_release_coalesce():
    a = FirstFit(100)
    x = a.alloc(30)
    y = a.alloc(50)
    z = a.alloc(30)
    assert (x, y, z) == (0, 30, None)
    a.release(x, 30)
    assert a.alloc(25) == 0
    a.release(0, 25)
    a.release(y, 50)
    assert a.free == [(0, 100)] and a.used == 0
    assert a.alloc(100) == 0 and a.alloc(1) is None


def test_budget_accounting_never_exceeds_capacity():
    rng = np.random.default_rng(0)
    a = FirstFit(10_000)
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.5:
            off, n = live.pop(rng.integers(len(live)))
            a.release(off, n)
        else:
            n = int(rng.integers(1, 900))
            off = a.alloc(n)
            if off is not None:
                live.append((off, n))
        assert a.used == sum(n for _, n in live) <= a.cap
        spans = sorted(live)
        for (o1, n1), (o2, _) in zip(spans, spans[1:]):
            assert o1 + n1 <= o2  # no overlap


def test_period_grid_scaling():
    g = period_grid(2048 * 0.0204, n=2048)
    assert len(g) == 500 and abs(g[0] - 0.5) < 1e-12 and abs(g[-1] - 2048 * 0.0204 / 3) < 1e-9
    assert len(period_grid(70000 * 0.0204, n=70000)) == 4000
    r = Request(1, np.zeros(10, np.float32), np.zeros(10, np.float32), np.zeros(3))
    assert r.bytes() == 10 * 12 + 3 * 36
