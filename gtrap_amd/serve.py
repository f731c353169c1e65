"""Served transit search: continuous batching of TLS+detrend requests on one GPU.

A request is one light curve (t, flux) with its trial-period grid.  Its result
is the best period of the periodic trapezoid search (gtrap_amd.tlsfold) run on
the median-detrended flux (flux / median, gtrap_amd.gfilter), with the epoch,
W, L, H and S/N of the best trial and the periodogram SDE.

Two paths compute exactly the same per-request result:

* :func:`search_plain` -- one launch sequence per curve: upload, detrend,
  search, statistics, download, synchronise (the reference's calling pattern,
  kepler/gtls_mockkepler.py:229-318, with the search swapped for tls_fold).
* :class:`TransitServer` -- a device-resident loop.  Requests are admitted into
  slots of preallocated HBM arenas (sample arena for t / flux / detrended
  flux, period arena for the grid and per-period results) as memory frees up
  (first-fit, budgeted in bytes).  Every step detrends the curves admitted since
  the previous step (ragged median filter), then scores a slice of
  (curve, period) work items drawn from all resident requests, oldest first,
  with the persistent tls_fold kernel (items handed out by an atomic ticket,
  longest first), then reduces the periodogram of every request whose last
  items ran in this step.  The step is a fixed-shape sequence of kernels whose
  variable parts (segment list, item list, counts) live in device descriptor
  buffers, so it is captured once as HIP graphs (detrend, fold, statistics)
  and replayed.  Requests that arrive while a step runs join the next one:
  their work overlaps with the work of requests admitted earlier.

Each work item's computation depends only on its own curve and period, so a
result is the same alone and in any mix (bit for bit).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .tlsfold import Workspace

SAMPLE_BYTES = 12   # t, flux, detrended flux (float32 each)
PERIOD_BYTES = 36   # period f64, S/N f64, H f64, w/l/e int32


@dataclass
class Request:
    rid: int
    t: np.ndarray
    flux: np.ndarray
    periods: np.ndarray
    arrival: float = 0.0

    def bytes(self) -> int:
        return SAMPLE_BYTES * len(self.t) + PERIOD_BYTES * len(self.periods)


@dataclass
class Result:
    rid: int
    best_period: float
    best_index: int
    sn: float
    sde: float
    w: int
    l: int
    e: int
    h: float
    arrival: float = 0.0
    done: float = 0.0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class FirstFit:
    """First-fit allocator over [0, cap) with coalescing frees."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        self.free = [(0, self.cap)]

    def alloc(self, size: int):
        size = int(size)
        for i, (a, n) in enumerate(self.free):
            if n >= size:
                if n == size:
                    self.free.pop(i)
                else:
                    self.free[i] = (a + size, n - size)
                return a
        return None

    def release(self, a: int, size: int):
        self.free.append((int(a), int(size)))
        self.free.sort()
        merged = []
        for s, n in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + n)
            else:
                merged.append((s, n))
        self.free = merged

    @property
    def used(self) -> int:
        return self.cap - sum(n for _, n in self.free)


# ----------------------------------------------------------------------------
# plain path: one launch sequence per curve
# ----------------------------------------------------------------------------
class PlainSearch:
    """Reusable buffers for :func:`search_plain` (the per-call allocations of
    the reference's drivers would otherwise dominate small requests)."""

    def __init__(self):
        self.ws = Workspace()


def search_plain(req: Request, widths, nl=10, dtep=0.0204, r=64, mode="div", nblocks=512,
                 ctx: PlainSearch | None = None) -> Result:
    import torch
    from . import gfilter
    from .tlsfold import FoldBatch, tlsfold_device
    _native.require_gpu()
    ctx = ctx or PlainSearch()
    n, m = len(req.t), len(req.periods)
    t = torch.from_numpy(np.ascontiguousarray(req.t, np.float32)).cuda()
    x = torch.from_numpy(np.ascontiguousarray(req.flux, np.float32)).cuda()
    per = torch.from_numpy(np.ascontiguousarray(req.periods, np.float64)).cuda()
    xd = torch.empty_like(x)
    gfilter.medfilt_device(x, xd, n, 1, r, mode)
    off = torch.zeros(1, dtype=torch.int64, device="cuda")
    ln = torch.full((1,), n, dtype=torch.int32, device="cuda")
    poff = torch.tensor([0, m], dtype=torch.int64, device="cuda")
    batch = FoldBatch(t, xd, off, ln, per, poff, 1, m, n, np.array([0, m]))
    out = tlsfold_device(batch, widths, nl, dtep, nblocks=nblocks, workspace=ctx.ws)
    bp = torch.empty(1, dtype=torch.int64, device="cuda")
    bs = torch.empty(1, dtype=torch.float64, device="cuda")
    sde = torch.empty(1, dtype=torch.float64, device="cuda")
    cur = torch.zeros(1, dtype=torch.int32, device="cuda")
    one = torch.ones(1, dtype=torch.int32, device="cuda")
    npc = torch.full((1,), m, dtype=torch.int32, device="cuda")
    p = _native.ptr
    _native.call("gtrap_tlsstats", p(out["sn"]), p(poff), p(npc), p(cur), p(one), 1, p(bp), p(bs),
                 p(sde), _native.stream_handle())
    k = int(bp.item())  # synchronises
    res = Result(req.rid, float(req.periods[k]) if k >= 0 else float("nan"), k, float(bs.item()),
                 float(sde.item()), int(out["w"][k]) if k >= 0 else -1,
                 int(out["l"][k]) if k >= 0 else -1, int(out["e"][k]) if k >= 0 else -1,
                 float(out["h"][k]) if k >= 0 else float("nan"), req.arrival)
    return res


# ----------------------------------------------------------------------------
# served path
# ----------------------------------------------------------------------------
@dataclass
class _Slot:
    req: Request
    soff: int
    poff: int
    next_item: int = 0
    admitted: float = 0.0


@dataclass
class ServerConfig:
    widths: np.ndarray = field(default_factory=lambda: np.linspace(0.04, 0.3, 10))
    nl: int = 10
    dtep: float = 0.0204
    r: int = 64
    mode: str = "div"
    hbm_budget: int = 4 << 30        # bytes for the sample and period arenas
    period_share: float = 0.25       # fraction of the budget for the period arena
    max_curves: int = 4096           # resident request slots
    max_len: int = 70000             # longest admissible curve
    items_per_step: int = 32768      # (curve, period) work items per step
    nblocks: int = 1024              # persistent tls_fold workgroups
    seg_len: int = 512               # detrend rows per lane
    use_graph: bool = True

    @classmethod
    def tuned(cls, **kw) -> "ServerConfig":
        """Defaults overridden by the autotuner's cached choices (if any)."""
        from . import autotune
        cfg = cls(**kw)
        if "items_per_step" not in kw:
            cfg.items_per_step = int(autotune.lookup("serve.items_per_step", cfg.items_per_step))
        if "nblocks" not in kw:
            cfg.nblocks = int(autotune.lookup("serve.nblocks", cfg.nblocks))
        return cfg


class TransitServer:
    def __init__(self, cfg: ServerConfig | None = None):
        import torch
        _native.require_gpu()
        self.cfg = cfg = cfg or ServerConfig()
        dev = "cuda"
        pbytes = int(cfg.hbm_budget * cfg.period_share)
        self.scap = max(cfg.max_len, (cfg.hbm_budget - pbytes) // SAMPLE_BYTES)
        self.pcap = max(1, pbytes // PERIOD_BYTES)
        self.salloc = FirstFit(self.scap)
        self.palloc = FirstFit(self.pcap)
        f32, f64, i32, i64 = torch.float32, torch.float64, torch.int32, torch.int64
        self.t = torch.zeros(self.scap, dtype=f32, device=dev)
        self.x = torch.zeros(self.scap, dtype=f32, device=dev)
        self.xd = torch.zeros(self.scap, dtype=f32, device=dev)
        self.per = torch.zeros(self.pcap, dtype=f64, device=dev)
        self.sn = torch.zeros(self.pcap, dtype=f64, device=dev)
        self.h = torch.zeros(self.pcap, dtype=f64, device=dev)
        self.bw = torch.zeros(self.pcap, dtype=i32, device=dev)
        self.bl = torch.zeros(self.pcap, dtype=i32, device=dev)
        self.be = torch.zeros(self.pcap, dtype=i32, device=dev)
        C = cfg.max_curves
        self.max_segs = C * ((cfg.max_len + cfg.seg_len - 1) // cfg.seg_len)
        # device descriptors (fixed addresses, captured by the graph)
        self.d_off = torch.zeros(C, dtype=i64, device=dev)
        self.d_len = torch.zeros(C, dtype=i32, device=dev)
        self.d_poff = torch.zeros(C, dtype=i64, device=dev)
        self.d_np = torch.zeros(C, dtype=i32, device=dev)
        self.d_seg_curve = torch.zeros(self.max_segs, dtype=i32, device=dev)
        self.d_seg_start = torch.zeros(self.max_segs, dtype=i32, device=dev)
        self.d_item_curve = torch.zeros(cfg.items_per_step, dtype=i32, device=dev)
        self.d_item_pidx = torch.zeros(cfg.items_per_step, dtype=i64, device=dev)
        self.d_stat_curves = torch.zeros(C, dtype=i32, device=dev)
        self.d_counts = torch.zeros(4, dtype=i32, device=dev)  # nseg, nitems, nstats, ticket
        self.d_best = torch.zeros(C, dtype=i64, device=dev)
        self.d_bsn = torch.zeros(C, dtype=f64, device=dev)
        self.d_sde = torch.zeros(C, dtype=f64, device=dev)
        self.widths = torch.from_numpy(np.ascontiguousarray(cfg.widths, np.float64)).to(dev)
        self.ws = Workspace()
        self.ws_buf, self.ws_bytes = self.ws.get(cfg.max_len, cfg.nblocks)
        # pinned host mirrors of the per-step descriptors
        pin = dict(pin_memory=True)
        self.h_off = torch.zeros(C, dtype=i64, **pin)
        self.h_len = torch.zeros(C, dtype=i32, **pin)
        self.h_poff = torch.zeros(C, dtype=i64, **pin)
        self.h_np = torch.zeros(C, dtype=i32, **pin)
        self.h_seg_curve = torch.zeros(self.max_segs, dtype=i32, **pin)
        self.h_seg_start = torch.zeros(self.max_segs, dtype=i32, **pin)
        self.h_item_curve = torch.zeros(cfg.items_per_step, dtype=i32, **pin)
        self.h_item_pidx = torch.zeros(cfg.items_per_step, dtype=i64, **pin)
        self.h_stat_curves = torch.zeros(C, dtype=i32, **pin)
        self.h_counts = torch.zeros(4, dtype=i32, **pin)
        self.free_slots = list(range(C - 1, -1, -1))
        self.slots: dict[int, _Slot] = {}
        self.order: list[int] = []   # resident slots, oldest first
        self.queue: list[Request] = []
        self.graphs = None
        self.fold_log = None  # list of (start event, end event, trials) when set
        self.steps = 0
        self.stream = torch.cuda.current_stream()

    # -- admission -----------------------------------------------------------
    def submit(self, req: Request) -> None:
        if len(req.t) > self.cfg.max_len:
            raise ValueError(f"request {req.rid}: {len(req.t)} samples > max_len")
        if len(req.t) > self.scap or len(req.periods) > self.pcap:
            raise ValueError(f"request {req.rid} exceeds the HBM budget")
        self.queue.append(req)

    def _admit(self, now: float):
        new = []
        while self.queue and self.free_slots and self.queue[0].arrival <= now:
            req = self.queue[0]
            so = self.salloc.alloc(len(req.t))
            if so is None:
                break
            po = self.palloc.alloc(len(req.periods))
            if po is None:
                self.salloc.release(so, len(req.t))
                break
            self.queue.pop(0)
            s = self.free_slots.pop()
            self.slots[s] = _Slot(req, so, po, 0, now)
            self.order.append(s)
            new.append(s)
        return new

    # -- one step ------------------------------------------------------------
    def _rec_pre(self):
        """Ticket reset and detrend of the newly admitted curves."""
        cfg = self.cfg
        p = _native.ptr
        cnt = self.d_counts
        cnt[3:4].zero_()  # ticket counter (a fill kernel, graph-safe)
        _native.call("gtrap_medfilt_ragged", p(self.x), p(self.xd), p(self.d_off), p(self.d_len),
                     p(self.d_seg_curve), p(self.d_seg_start), p(cnt[0:1]), self.max_segs,
                     cfg.r, 1 if cfg.mode == "div" else 0, cfg.seg_len, _native.stream_handle())

    def _rec_fold(self):
        """The step's (curve, period) work items."""
        cfg = self.cfg
        p = _native.ptr
        cnt = self.d_counts
        _native.call("gtrap_tlsfold_items", p(self.sn), p(self.h), p(self.bw), p(self.bl),
                     p(self.be), p(self.t), p(self.xd), p(self.d_off), p(self.d_len),
                     cfg.max_len, p(self.per), p(self.d_item_curve), p(self.d_item_pidx),
                     p(cnt[1:2]), p(cnt[3:4]), p(self.widths), int(self.widths.numel()), cfg.nl,
                     float(cfg.dtep), p(self.ws_buf), int(self.ws_bytes), cfg.nblocks,
                     _native.stream_handle())

    def _rec_post(self):
        """Periodogram statistics of the curves finished in this step."""
        cfg = self.cfg
        p = _native.ptr
        _native.call("gtrap_tlsstats", p(self.sn), p(self.d_poff), p(self.d_np),
                     p(self.d_stat_curves), p(self.d_counts[2:3]), cfg.max_curves,
                     p(self.d_best), p(self.d_bsn), p(self.d_sde), _native.stream_handle())

    def _launch(self, trials: int):
        """One step's fixed kernel sequence: three graphs (detrend, fold,
        statistics) captured on the first step and replayed after it.  With
        ``fold_log`` set, HIP events on the serving stream bracket the fold."""
        import torch
        parts = (self._rec_pre, self._rec_fold, self._rec_post)
        if self.cfg.use_graph and self.graphs is None:
            # the capture does not execute the kernels: this step runs eagerly
            torch.cuda.synchronize()
            gs = []
            for f in parts:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    f()
                gs.append(g)
            self.graphs = gs
            run = parts
        elif self.cfg.use_graph:
            run = [g.replay for g in self.graphs]
        else:
            run = parts
        run[0]()
        if self.fold_log is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record(self.stream)
            run[1]()
            ev[1].record(self.stream)
            self.fold_log.append((ev[0], ev[1], trials))
        else:
            run[1]()
        run[2]()

    def fold_stats(self):
        """(launches, mean fold kernel ms, total trials, total fold seconds) of
        the logged steps; synchronises."""
        import torch
        torch.cuda.synchronize()
        ms = [a.elapsed_time(b) for a, b, _ in self.fold_log]
        tr = sum(k for _, _, k in self.fold_log)
        return len(ms), (float(np.mean(ms)) if ms else 0.0), tr, sum(ms) * 1e-3

    def step(self, now: float | None = None) -> list[Result]:
        """Admit arrived requests, run one device step, return finished results."""
        import torch
        cfg = self.cfg
        now = time.perf_counter() if now is None else now
        new = self._admit(now)
        # uploads of newly admitted curves (pinned staging through torch)
        for s in new:
            sl = self.slots[s]
            n, m = len(sl.req.t), len(sl.req.periods)
            self.t[sl.soff:sl.soff + n].copy_(torch.from_numpy(
                np.ascontiguousarray(sl.req.t, np.float32)), non_blocking=False)
            self.x[sl.soff:sl.soff + n].copy_(torch.from_numpy(
                np.ascontiguousarray(sl.req.flux, np.float32)), non_blocking=False)
            self.per[sl.poff:sl.poff + m].copy_(torch.from_numpy(
                np.ascontiguousarray(sl.req.periods, np.float64)), non_blocking=False)
            self.h_off[s] = sl.soff
            self.h_len[s] = n
            self.h_poff[s] = sl.poff
            self.h_np[s] = m
        # detrend segments of the new curves
        ns = 0
        for s in new:
            n = len(self.slots[s].req.t)
            k = (n + cfg.seg_len - 1) // cfg.seg_len
            self.h_seg_curve[ns:ns + k] = s
            self.h_seg_start[ns:ns + k] = torch.arange(0, k * cfg.seg_len, cfg.seg_len,
                                                       dtype=torch.int32)
            ns += k
        # work items: oldest requests first, up to items_per_step
        budget = cfg.items_per_step
        ic, ip, cost, done = [], [], [], []
        nep = 0
        for s in self.order:
            if budget == 0:
                break
            sl = self.slots[s]
            m = len(sl.req.periods)
            take = min(m - sl.next_item, budget)
            if take <= 0:
                continue
            a = sl.next_item
            per = sl.req.periods[a:a + take]
            ic.append(np.full(take, s, np.int32))
            ip.append(np.arange(sl.poff + a, sl.poff + a + take, dtype=np.int64))
            cost.append(len(sl.req.t) + per * (100.0 / cfg.dtep))
            nep += int(np.ceil(per / cfg.dtep).sum())
            sl.next_item += take
            budget -= take
            if sl.next_item == m:
                done.append(s)
        ni = cfg.items_per_step - budget
        if ni:
            ic, ip, cost = np.concatenate(ic), np.concatenate(ip), np.concatenate(cost)
            o = np.argsort(-cost, kind="stable")  # longest items first
            self.h_item_curve[:ni] = torch.from_numpy(ic[o])
            self.h_item_pidx[:ni] = torch.from_numpy(ip[o])
        for k, s in enumerate(done):
            self.h_stat_curves[k] = s
        self.h_counts[0], self.h_counts[1], self.h_counts[2] = ns, ni, len(done)
        # descriptors -> device (same stream, before the step's kernels)
        for d, h in ((self.d_off, self.h_off), (self.d_len, self.h_len),
                     (self.d_poff, self.h_poff), (self.d_np, self.h_np)):
            d.copy_(h, non_blocking=True)
        if ns:
            self.d_seg_curve[:ns].copy_(self.h_seg_curve[:ns], non_blocking=True)
            self.d_seg_start[:ns].copy_(self.h_seg_start[:ns], non_blocking=True)
        if ni:
            self.d_item_curve[:ni].copy_(self.h_item_curve[:ni], non_blocking=True)
            self.d_item_pidx[:ni].copy_(self.h_item_pidx[:ni], non_blocking=True)
        if done:
            self.d_stat_curves[:len(done)].copy_(self.h_stat_curves[:len(done)],
                                                 non_blocking=True)
        self.d_counts[:3].copy_(self.h_counts[:3], non_blocking=True)
        self._launch(nep * len(cfg.widths) * cfg.nl)
        self.steps += 1
        if not done:
            torch.cuda.current_stream().synchronize()
            return []
        idx = torch.tensor(done, dtype=torch.int64, device="cuda")
        best = self.d_best[idx].cpu().numpy()
        bsn = self.d_bsn[idx].cpu().numpy()
        sde = self.d_sde[idx].cpu().numpy()
        bi = torch.from_numpy(np.maximum(best, 0)).cuda()
        w = self.bw[bi].cpu().numpy()
        l_ = self.bl[bi].cpu().numpy()
        e = self.be[bi].cpu().numpy()
        h = self.h[bi].cpu().numpy()
        tdone = time.perf_counter()
        out = []
        for k, s in enumerate(done):
            sl = self.slots.pop(s)
            self.order.remove(s)
            j = int(best[k]) - sl.poff if best[k] >= 0 else -1
            out.append(Result(sl.req.rid, float(sl.req.periods[j]) if j >= 0 else float("nan"),
                              j, float(bsn[k]), float(sde[k]),
                              int(w[k]) if j >= 0 else -1, int(l_[k]) if j >= 0 else -1,
                              int(e[k]) if j >= 0 else -1, float(h[k]) if j >= 0 else float("nan"),
                              sl.req.arrival, tdone))
            self.salloc.release(sl.soff, len(sl.req.t))
            self.palloc.release(sl.poff, len(sl.req.periods))
            self.free_slots.append(s)
        return out

    def busy(self) -> bool:
        return bool(self.queue or self.slots)

    def run(self, requests, t0: float | None = None, realtime: bool = True) -> list[Result]:
        """Serve a trace.  ``arrival`` of each request is seconds after t0; with
        realtime=False every request is available immediately."""
        t0 = time.perf_counter() if t0 is None else t0
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.arrival = t0 + (r.arrival if realtime else 0.0)
            self.submit(r)
        results = []
        while self.busy():
            now = time.perf_counter()
            if self.queue and not self.slots and self.queue[0].arrival > now:
                time.sleep(min(1e-3, self.queue[0].arrival - now))
                continue
            results.extend(self.step(now))
        return results
