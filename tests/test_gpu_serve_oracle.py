"""The served TLS+detrend path against the oracle chain at serving sizes.

Requests of 2048 to 70000 samples with their full (unthinned) period grids are
served through TransitServer with captured graphs.  Every request is checked:

* served == plain, bit for bit, for the result and the whole periodogram;
* the ragged detrend: the served S/N of sampled periods equals the oracle
  chain scipy.signal.medfilt (r = 64, divide) -> direct tls_fold oracle at
  1e-7 relative, best (W, L, epoch) identical except near-ties;
* the periodogram statistics: best period, best S/N and SDE equal the oracle's
  tlsstats over the served periodogram (SDE 1e-9 relative);
* the best trial itself re-evaluated by the oracle's single-trial function.

Detrend of ragged segments is also checked on its own against scipy, bit for
bit.
"""
import sys

import numpy as np
import pytest

from oracle import medfilt as om
from oracle import tlsfold as ot

pytestmark = pytest.mark.gpu

LENGTHS = (2048, 3100, 5200, 9000, 17000, 31000, 45000, 70000)
WIDTHS = np.linspace(0.04, 0.3, 10)
NL, DTEP = 10, 0.0204


def _requests():
    from gtrap_amd import synth
    from gtrap_amd.serve import Request
    from gtrap_amd.tlsfold import period_grid
    rng = np.random.default_rng(4040)
    out = []
    for i, n in enumerate(LENGTHS):
        t = np.arange(n) * DTEP
        base = n * DTEP
        flux = 1.0 + 2e-3 * np.sin(2 * np.pi * t / rng.uniform(5, 50)) + \
            1e-3 * rng.standard_normal(n)
        if i % 2 == 0:
            P = rng.uniform(0.5, base / 3)
            t14 = rng.uniform(0.05, 0.3)
            flux -= synth.trapezoid_dip(t, rng.uniform(0, P), P, rng.uniform(1e-3, 5e-3), t14,
                                        t14 / 10)
        out.append(Request(i, t.astype(np.float32), flux.astype(np.float32),
                           period_grid(base, n=n), arrival=0.0))
    return out


def test_ragged_detrend_bit_exact(gpu):
    import torch
    from gtrap_amd import _native
    rng = np.random.default_rng(5)
    lens = [7, 64, 129, 1000, 5000, 20011]
    for r in (7, 24, 64, 100, 128, 512, 700, 1500):
        xs = [(1 + 1e-3 * rng.standard_normal(n)).astype(np.float32) for n in lens]
        off = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.int64)
        flat = np.concatenate(xs)
        seg = 512
        sc, ss = [], []
        for c, n in enumerate(lens):
            for s0 in range(0, n, seg):
                sc.append(c)
                ss.append(s0)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        x, y = dev(flat), torch.zeros(len(flat), dtype=torch.float32, device="cuda")
        p = _native.ptr
        nseg = dev(np.array([len(sc)], np.int32))
        d_off, d_len = dev(off), dev(np.array(lens, np.int32))
        d_sc, d_ss = dev(np.array(sc, np.int32)), dev(np.array(ss, np.int32))
        _native.call("gtrap_medfilt_ragged", p(x), p(y), p(d_off), p(d_len), p(d_sc), p(d_ss),
                     p(nseg), len(sc) + 3, r, 1, seg, _native.stream_handle())
        got = y.cpu().numpy()
        for c, n in enumerate(lens):
            want = om.detrend_scipy(xs[c][:, None], r, "div")[:, 0]
            np.testing.assert_array_equal(got[off[c]:off[c] + n].view(np.uint32),
                                          want.view(np.uint32), err_msg=f"r={r} curve {c}")


def test_served_matches_plain_and_oracle_at_serving_sizes(gpu):
    from gtrap_amd.serve import ServerConfig, TransitServer, search_plain
    reqs = _requests()
    cfg = ServerConfig(items_per_step=3000, max_curves=16, max_len=70000,
                       hbm_budget=256 << 20, nblocks=512, use_graph=True, periodogram=True)
    srv = TransitServer(cfg)
    print("serving", [len(r.t) for r in reqs], flush=True)
    got = {r.rid: r for r in srv.run(_requests(), realtime=False)}
    assert len(got) == len(reqs) and all(g.ok for g in got.values()) and srv.steps > 3
    rng = np.random.default_rng(7)
    for req in reqs:
        g = got[req.rid]
        n, m = len(req.t), len(req.periods)
        print(f"request {req.rid}: n={n} periods={m}", flush=True)
        # served == plain, bit for bit
        pl = search_plain(req, cfg.widths, cfg.nl, cfg.dtep, cfg.r, periodogram=True)
        assert (g.best_index, g.w, g.l, g.e) == (pl.best_index, pl.w, pl.l, pl.e)
        assert g.sn == pl.sn and g.h == pl.h and g.sde == pl.sde
        for k in ("sn", "h", "w", "l", "e"):
            np.testing.assert_array_equal(g.periodogram[k], pl.periodogram[k])
        pg = g.periodogram
        # statistics of the served periodogram
        best, bval, sde = ot.tlsstats(pg["sn"])
        assert best == g.best_index and bval == g.sn
        np.testing.assert_allclose(g.sde, sde, rtol=1e-9)
        # sampled periods against the oracle chain (scipy detrend -> direct search)
        xd = om.detrend_scipy(req.flux[:, None], cfg.r, "div")[:, 0]
        ns = 10 if n < 30000 else 5
        idx = np.unique(np.concatenate([[0, m - 1, g.best_index],
                                        rng.choice(m, ns, replace=False)]))
        want = ot.tlsfold(req.t, xd, req.periods[idx], WIDTHS, NL, DTEP)
        fin = np.isfinite(want["sn"])
        np.testing.assert_array_equal(np.isfinite(pg["sn"][idx]), fin)
        np.testing.assert_allclose(pg["sn"][idx][fin], want["sn"][fin], rtol=1e-7)
        for j, p in enumerate(idx):
            if (pg["w"][p], pg["l"][p], pg["e"][p]) != (want["w"][j], want["l"][j], want["e"][j]):
                s = ot.trial(req.t, xd, req.periods[p], pg["e"][p], WIDTHS[pg["w"][p]],
                             pg["l"][p], NL, DTEP)[0]
                assert abs(s - want["sn"][j]) <= 1e-9 * abs(want["sn"][j]), (req.rid, p)
        # the best trial re-evaluated directly
        s, H, _ = ot.trial(req.t, xd, g.best_period, g.e, WIDTHS[g.w], g.l, NL, DTEP)
        np.testing.assert_allclose([g.sn, g.h], [s, H], rtol=1e-7)
        sys.stdout.flush()
