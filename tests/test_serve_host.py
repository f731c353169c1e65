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


# ----------------------------------------------------------------------------
# serving loop against a CPU backend: scheduling, failure isolation, cancelling
# ----------------------------------------------------------------------------
import pytest

from gtrap_amd.serve import ServerConfig, TransitServer, SAMPLE_BYTES, PERIOD_BYTES
from oracle.tlsfold import tlsstats


def _fake_sn(req, j):
    """Deterministic per-period S/N of a request (a stand-in for the search)."""
    return float(np.cos(req.periods[j] * (req.rid + 1.7)) + 1e-3 * req.flux[j % len(req.flux)])


class FakeBackend:
    """The device side of TransitServer on the CPU: every launched work item
    gets _fake_sn, and the statistics of finished requests come from the
    oracle's tlsstats.  Failures are injected by request id."""

    def __init__(self, cfg, fail_upload=(), fail_step=()):
        pbytes = int(cfg.hbm_budget * cfg.period_share)
        self.scap = max(cfg.max_len, (cfg.hbm_budget - pbytes) // SAMPLE_BYTES)
        self.pcap = max(1, pbytes // PERIOD_BYTES)
        self.cfg = cfg
        self.fail_upload, self.fail_step = set(fail_upload), set(fail_step)
        self.req = {}
        self.sn = {}
        self.fold_log = None
        self.launches = 0

    def upload(self, slot, soff, poff, req):
        if req.rid in self.fail_upload:
            raise RuntimeError("injected upload failure")
        self.req[slot] = (req, poff)

    def release(self, slot):
        self.req.pop(slot, None)

    def launch(self, plan):
        for s in set(plan.item_curve.tolist()):
            if self.req[s][0].rid in self.fail_step:
                self.fail_step.discard(self.req[s][0].rid)
                raise RuntimeError("injected launch failure")
        self.launches += 1
        for s, p in zip(plan.item_curve.tolist(), plan.item_pidx.tolist()):
            req, poff = self.req[s]
            self.sn[p] = _fake_sn(req, p - poff)
        out = {}
        for s in plan.done:
            req, poff = self.req[s]
            v = np.array([self.sn[poff + j] for j in range(len(req.periods))])
            best, val, sde = tlsstats(v)
            out[s] = (poff + best, val, sde, 1, 2, 3, 0.5, None)
        return out

    def collect(self, handle):
        return handle


def _reqs(k=12, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(rng.integers(200, 3000))
        m = int(rng.integers(5, 400))
        out.append(Request(i, np.arange(n, dtype=np.float32) * 0.02,
                           (1 + 1e-3 * rng.standard_normal(n)).astype(np.float32),
                           np.sort(rng.uniform(0.5, 20, m)), arrival=i * 1e-4))
    return out


def _alone(req):
    v = np.array([_fake_sn(req, j) for j in range(len(req.periods))])
    return tlsstats(v)


def _cfg(**kw):
    base = dict(items_per_step=97, max_curves=6, max_len=3000, hbm_budget=1 << 20)
    base.update(kw)
    return ServerConfig(**base)


@pytest.mark.parametrize("items", [13, 97, 5000])
def test_every_request_gets_its_own_result(items):
    reqs = _reqs()
    srv = TransitServer(_cfg(items_per_step=items), backend=FakeBackend(_cfg()))
    res = srv.run(reqs, realtime=False)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    for r in res:
        best, val, sde = _alone(reqs[r.rid])
        assert r.ok and r.best_index == best and r.sn == val and r.sde == sde
        assert r.best_period == reqs[r.rid].periods[best]
    assert srv.salloc.used == 0 and srv.palloc.used == 0 and not srv.slots


def test_upload_failure_retires_only_that_request():
    reqs = _reqs()
    srv = TransitServer(_cfg(), backend=FakeBackend(_cfg(), fail_upload={3, 7}))
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert set(res) == set(range(len(reqs)))
    for rid, r in res.items():
        if rid in (3, 7):
            assert not r.ok and "upload" in r.error
        else:
            assert r.ok and r.best_index == _alone(reqs[rid])[0]
    assert srv.salloc.used == 0 and srv.palloc.used == 0


def test_step_failure_retires_the_requests_in_that_step_only():
    reqs = _reqs(k=16, seed=1)
    srv = TransitServer(_cfg(items_per_step=60), backend=FakeBackend(_cfg(), fail_step={5}))
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert set(res) == set(range(16))
    failed = {rid for rid, r in res.items() if not r.ok}
    assert 5 in failed and len(failed) < 16
    for rid, r in res.items():
        if r.ok:
            best, val, sde = _alone(reqs[rid])
            assert r.best_index == best and r.sn == val
    assert srv.salloc.used == 0 and srv.palloc.used == 0 and not srv.slots


def test_cancel_queued_and_resident():
    reqs = _reqs(k=10, seed=2)
    cfg = _cfg(items_per_step=40, max_curves=3)
    srv = TransitServer(cfg, backend=FakeBackend(cfg))
    for r in reqs:
        r.arrival = 0.0
        srv.submit(r)
    out = srv.step(now=1.0)
    out += srv.step(now=1.0)
    resident = [sl.req.rid for sl in srv.slots.values()]
    assert resident and srv.queue
    assert srv.cancel(resident[0]) and srv.cancel(srv.queue[-1].rid)
    assert not srv.cancel(12345)
    while srv.busy():
        out += srv.step(now=1.0)
    out += srv.flush()
    got = {r.rid for r in out}
    assert got == set(range(10)) - {resident[0], reqs[-1].rid}
    assert all(r.ok for r in out)
    assert srv.salloc.used == 0 and srv.palloc.used == 0 and len(srv.free_slots) == 3


def test_submit_validation():
    cfg = _cfg()
    srv = TransitServer(cfg, backend=FakeBackend(cfg))
    t = np.arange(100, dtype=np.float32)
    f = np.ones(100, np.float32)
    with pytest.raises(ValueError, match="empty period grid"):
        srv.submit(Request(1, t, f, np.zeros(0)))
    bad = f.copy()
    bad[7] = np.nan
    with pytest.raises(ValueError, match="finite"):
        srv.submit(Request(2, t, bad, np.array([1.0])))
    with pytest.raises(ValueError, match="max_len"):
        srv.submit(Request(3, np.zeros(4000, np.float32), np.ones(4000, np.float32),
                           np.array([1.0])))
    with pytest.raises(ValueError, match="positive"):
        srv.submit(Request(4, t, f, np.array([1.0, -2.0])))
    assert not srv.queue
