"""Served (continuous batching) vs plain (one launch sequence per curve).

Each request's result must be identical alone and in the mix: every work
item is computed from its own curve and period only.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(k=10, seed=0, nmin=2048, nmax=9000):
    from gtrap_amd import synth
    from gtrap_amd.serve import Request
    from gtrap_amd.tlsfold import period_grid
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        t = np.arange(n) * 0.0204
        base = n * 0.0204
        flux = 1.0 + 1e-3 * rng.standard_normal(n) + 2e-3 * np.sin(2 * np.pi * t / 30.0)
        if rng.random() < 0.5:
            P = rng.uniform(0.6, base / 3)
            flux -= synth.trapezoid_dip(t, rng.uniform(0, P), P, 3e-3, 0.15, 0.015)
        per = period_grid(base, n=n)[::8]  # thinned grid keeps the test short
        out.append(Request(i, t.astype(np.float32), flux.astype(np.float32), per,
                           arrival=float(i) * 1e-3))
    return out


def _same(a, b):
    assert a.rid == b.rid
    assert a.best_index == b.best_index and (a.w, a.l, a.e) == (b.w, b.l, b.e)
    assert a.sn == b.sn and a.h == b.h
    assert a.sde == b.sde or (np.isnan(a.sde) and np.isnan(b.sde))


@pytest.mark.parametrize("use_graph", [False, True])
def test_served_equals_plain(gpu, use_graph):
    from gtrap_amd.serve import ServerConfig, TransitServer, search_plain
    reqs = _requests()
    cfg = ServerConfig(items_per_step=97, max_curves=8, max_len=9000, hbm_budget=64 << 20,
                       nblocks=256, use_graph=use_graph)
    plain = {r.rid: search_plain(r, cfg.widths, cfg.nl, cfg.dtep, cfg.r) for r in reqs}
    srv = TransitServer(cfg)
    got = srv.run([r for r in _requests()], realtime=False)
    assert len(got) == len(reqs) and srv.steps > 3
    for g in got:
        _same(g, plain[g.rid])


def test_admission_by_budget(gpu):
    """A budget that holds ~2 requests: the rest wait for memory and still finish."""
    from gtrap_amd.serve import ServerConfig, TransitServer
    reqs = _requests(k=6, seed=1, nmin=3000, nmax=4000)
    need = max(r.bytes() for r in reqs)
    cfg = ServerConfig(items_per_step=300, max_curves=8, max_len=4000, hbm_budget=3 * need,
                       period_share=0.5, nblocks=128)
    srv = TransitServer(cfg)
    peak = 0
    for r in reqs:
        r.arrival = 0.0
        srv.submit(r)
    res = []
    while srv.busy():
        res.extend(srv.step(now=1e30))
        peak = max(peak, len(srv.slots))
        assert srv.salloc.used <= srv.scap and srv.palloc.used <= srv.pcap
    assert len(res) == 6 and peak < 6


def test_served_equals_plain_wide_detrend(gpu):
    """The drivers' wide windows (r = 128, kepler/gtls_mockkepler.py:229) in the
    ragged detrend of the serving loop."""
    from gtrap_amd.serve import ServerConfig, TransitServer, search_plain
    reqs = _requests(k=5, seed=3)
    cfg = ServerConfig(items_per_step=131, max_curves=8, max_len=9000, hbm_budget=64 << 20,
                       nblocks=256, r=128)
    plain = {r.rid: search_plain(r, cfg.widths, cfg.nl, cfg.dtep, cfg.r) for r in reqs}
    got = TransitServer(cfg).run(_requests(k=5, seed=3), realtime=False)
    assert len(got) == 5
    for g in got:
        _same(g, plain[g.rid])
