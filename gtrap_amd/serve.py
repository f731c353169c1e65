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
  and replayed.

Steps are pipelined: while step k runs on the device, the host admits the
requests that arrived, stages their data through pinned buffers and uploads
it on a side stream, and plans step k+1; the results of step k are gathered
on the device and copied to pinned host memory behind its kernels, so
collecting them waits for step k only.  Requests that arrive while a step runs
join the next one: their work overlaps with the work of requests admitted
earlier.

Each work item's computation depends only on its own curve and period, so a
result is the same alone and in any mix (bit for bit).

Failure handling: a request that fails validation is refused by
:meth:`TransitServer.submit`; a request whose upload fails is retired alone
with an error result; a step whose launch fails retires the requests that had
work items in it (error results) and the server keeps serving the others;
:meth:`TransitServer.cancel` removes a queued or resident request between
steps.  The device side is a backend object (:class:`HipBackend`), so the
host logic runs on the CPU against a test backend.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .tlsfold import STATS, Workspace

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
    error: str | None = None
    periodogram: dict | None = None   # per-period sn, h, w, l, e (ServerConfig.periodogram)

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def ok(self) -> bool:
        return self.error is None


def _failed(req: Request, msg: str, now: float) -> Result:
    return Result(req.rid, float("nan"), -1, float("nan"), float("nan"), -1, -1, -1,
                  float("nan"), req.arrival, now, error=msg)


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


def validate(req: Request, max_len: int | None = None) -> None:
    """Raise ValueError for a request the search cannot take: no periods, a
    non-positive or non-finite period, mismatched or empty arrays, non-finite
    flux (the median detrend assumes finite values), or more than max_len
    samples."""
    n = len(req.t)
    if n == 0 or len(req.flux) != n:
        raise ValueError(f"request {req.rid}: t and flux must be non-empty and of equal length")
    if max_len is not None and n > max_len:
        raise ValueError(f"request {req.rid}: {n} samples > max_len {max_len}")
    per = np.asarray(req.periods, np.float64)
    if per.size == 0:
        raise ValueError(f"request {req.rid}: empty period grid")
    if not np.all(np.isfinite(per) & (per > 0)):
        raise ValueError(f"request {req.rid}: periods must be finite and positive")
    if not np.all(np.isfinite(np.asarray(req.flux, np.float32))):
        raise ValueError(f"request {req.rid}: flux must be finite (mask samples by a "
                         "negative time instead)")


# ----------------------------------------------------------------------------
# plain path: one launch sequence per curve
# ----------------------------------------------------------------------------
class PlainSearch:
    """Reusable buffers for :func:`search_plain` (the per-call allocations of
    the reference's drivers would otherwise dominate small requests)."""

    def __init__(self):
        self.ws = Workspace()


def search_plain(req: Request, widths, nl=10, dtep=0.0204, r=64, mode="div", nblocks=512,
                 ctx: PlainSearch | None = None, periodogram: bool = False,
                 stat: str = "snr") -> Result:
    import torch
    from . import gfilter
    from .tlsfold import FoldBatch, tlsfold_device
    _native.require_gpu()
    validate(req)
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
    out = tlsfold_device(batch, widths, nl, dtep, nblocks=nblocks, workspace=ctx.ws, stat=stat)
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
    if periodogram:
        res.periodogram = {key: v.cpu().numpy() for key, v in out.items()}
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
    periodogram: bool = False        # return every request's per-period results
    stat: str = "snr"                # ranking statistic: "snr" (tls.py) or "dchi2"

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


@dataclass
class StepPlan:
    """Host description of one device step."""
    new: list            # slots admitted since the previous step (detrend)
    seg_curve: np.ndarray
    seg_start: np.ndarray
    item_curve: np.ndarray
    item_pidx: np.ndarray
    done: list           # slots whose last items run in this step (statistics)
    slot_off: dict       # slot -> (sample offset, length, period offset, period count)
    trials: int = 0


class _Pinned:
    """Grow-only pinned host buffer of one dtype."""

    def __init__(self, dtype):
        self.dtype = dtype
        self.buf = None

    def get(self, n: int):
        import torch
        if self.buf is None or self.buf.numel() < n:
            self.buf = torch.empty(max(n, 1 << 12), dtype=self.dtype, pin_memory=True)
        return self.buf[:n]


class HipBackend:
    """Device side of a :class:`TransitServer`: arenas, descriptors, the three
    captured step graphs, pinned staging for uploads (side stream) and for the
    results of a step (gathered on the device behind the step)."""

    def __init__(self, cfg: ServerConfig):
        import torch
        from . import autotune
        _native.require_gpu()
        autotune.apply_tiles()
        self.cfg = cfg
        dev = "cuda"
        pbytes = int(cfg.hbm_budget * cfg.period_share)
        self.scap = max(cfg.max_len, (cfg.hbm_budget - pbytes) // SAMPLE_BYTES)
        self.pcap = max(1, pbytes // PERIOD_BYTES)
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
        # device descriptors (fixed addresses, captured by the graphs)
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
        self.d_done = torch.zeros(C, dtype=i64, device=dev)
        self.widths = torch.from_numpy(np.ascontiguousarray(cfg.widths, np.float64)).to(dev)
        self.ws = Workspace()
        self.ws_buf, self.ws_bytes = self.ws.get(cfg.max_len, cfg.nblocks)
        # two sets of pinned host descriptors: a set is rewritten only after
        # the step that read it has completed
        pin = dict(pin_memory=True)
        self.hdesc = [dict(off=torch.zeros(C, dtype=i64, **pin), len=torch.zeros(C, dtype=i32, **pin),
                       poff=torch.zeros(C, dtype=i64, **pin), np=torch.zeros(C, dtype=i32, **pin),
                       seg_curve=torch.zeros(self.max_segs, dtype=i32, **pin),
                       seg_start=torch.zeros(self.max_segs, dtype=i32, **pin),
                       item_curve=torch.zeros(cfg.items_per_step, dtype=i32, **pin),
                       item_pidx=torch.zeros(cfg.items_per_step, dtype=i64, **pin),
                       stat=torch.zeros(C, dtype=i32, **pin), done=torch.zeros(C, dtype=i64, **pin),
                       counts=torch.zeros(4, dtype=i32, **pin))
                  for _ in range(2)]
        self.slot_tab = {}            # slot -> (soff, n, poff, m), mirrored into both sets
        self.stage = [(_Pinned(f32), _Pinned(f64)) for _ in range(2)]
        self.stage_ev = [None, None]  # upload of a staging set finished
        self.res_host = [_Pinned(f64) for _ in range(2)]
        self.pg_host = [[] for _ in range(2)]
        self.graphs = None
        self.fold_log = None          # list of (start event, end event, trials) when set
        self.stream = torch.cuda.current_stream()
        self.side = torch.cuda.Stream()   # default priority
        self.nsteps = 0
        self._pending_uploads = []    # (slot, soff, poff, req) staged for the next step
        self.side.wait_stream(self.stream)  # arenas are zeroed on the serving stream

    # -- uploads ---------------------------------------------------------------
    def upload(self, slot: int, soff: int, poff: int, req: Request) -> None:
        """Stage a newly admitted request; the copies are issued with the next
        step (pinned staging, side stream)."""
        self._pending_uploads.append((slot, soff, poff, req))
        self.slot_tab[slot] = (soff, len(req.t), poff, len(req.periods))

    def release(self, slot: int) -> None:
        self.slot_tab.pop(slot, None)

    def _issue_uploads(self, k: int):
        import torch
        ups = self._pending_uploads
        self._pending_uploads = []
        if not ups:
            return None
        if self.stage_ev[k] is not None:
            self.stage_ev[k].synchronize()   # staging set k free again
        nf = sum(2 * len(r.t) for _, _, _, r in ups)
        nd = sum(len(r.periods) for _, _, _, r in ups)
        sf = self.stage[k][0].get(nf).numpy()
        sd = self.stage[k][1].get(nd).numpy()
        a = b = 0
        spans = []
        for slot, soff, poff, r in ups:
            n, m = len(r.t), len(r.periods)
            sf[a:a + n] = r.t
            sf[a + n:a + 2 * n] = r.flux
            sd[b:b + m] = r.periods
            spans.append((soff, poff, n, m, a, b))
            a += 2 * n
            b += m
        tf, td = self.stage[k][0].buf, self.stage[k][1].buf
        with torch.cuda.stream(self.side):
            for soff, poff, n, m, a, b in spans:
                self.t[soff:soff + n].copy_(tf[a:a + n], non_blocking=True)
                self.x[soff:soff + n].copy_(tf[a + n:a + 2 * n], non_blocking=True)
                self.per[poff:poff + m].copy_(td[b:b + m], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.side)
        self.stage_ev[k] = ev
        return ev

    # -- one step --------------------------------------------------------------
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
                     float(cfg.dtep), STATS[cfg.stat], p(self.ws_buf), int(self.ws_bytes),
                     cfg.nblocks, _native.stream_handle())

    def _rec_post(self):
        """Periodogram statistics of the curves finished in this step."""
        cfg = self.cfg
        p = _native.ptr
        _native.call("gtrap_tlsstats", p(self.sn), p(self.d_poff), p(self.d_np),
                     p(self.d_stat_curves), p(self.d_counts[2:3]), cfg.max_curves,
                     p(self.d_best), p(self.d_bsn), p(self.d_sde), _native.stream_handle())

    def launch(self, plan: StepPlan):
        """Issue one step (asynchronous): descriptors, the three graphs, the
        device gather of the finished requests' results into pinned memory.
        Returns a handle for :meth:`collect`.  At most two steps may be in
        flight (the pinned descriptor, staging and result sets alternate)."""
        import torch
        k = self.nsteps & 1
        up = self._issue_uploads(k)
        H = self.hdesc[k]
        for s, (soff, n, poff, m) in plan.slot_off.items():
            H["off"][s], H["len"][s], H["poff"][s], H["np"][s] = soff, n, poff, m
        ns, ni, nd = len(plan.seg_curve), len(plan.item_curve), len(plan.done)
        if ns:
            H["seg_curve"][:ns] = torch.from_numpy(plan.seg_curve)
            H["seg_start"][:ns] = torch.from_numpy(plan.seg_start)
        if ni:
            H["item_curve"][:ni] = torch.from_numpy(plan.item_curve)
            H["item_pidx"][:ni] = torch.from_numpy(plan.item_pidx)
        for j, s in enumerate(plan.done):
            H["stat"][j] = s
            H["done"][j] = s
        H["counts"][0], H["counts"][1], H["counts"][2] = ns, ni, nd
        st = self.stream
        with torch.cuda.stream(st):
            if up is not None:
                st.wait_event(up)
            for d, h in ((self.d_off, H["off"]), (self.d_len, H["len"]),
                         (self.d_poff, H["poff"]), (self.d_np, H["np"])):
                d.copy_(h, non_blocking=True)
            if ns:
                self.d_seg_curve[:ns].copy_(H["seg_curve"][:ns], non_blocking=True)
                self.d_seg_start[:ns].copy_(H["seg_start"][:ns], non_blocking=True)
            if ni:
                self.d_item_curve[:ni].copy_(H["item_curve"][:ni], non_blocking=True)
                self.d_item_pidx[:ni].copy_(H["item_pidx"][:ni], non_blocking=True)
            if nd:
                self.d_stat_curves[:nd].copy_(H["stat"][:nd], non_blocking=True)
                self.d_done[:nd].copy_(H["done"][:nd], non_blocking=True)
            self.d_counts[:3].copy_(H["counts"][:3], non_blocking=True)
            self._run_graphs(plan.trials)
            spans = []
            if nd:
                # gather best, S/N, SDE and the best trial's w, l, e, H per request
                idx = self.d_done[:nd]
                best = self.d_best[idx]
                bi = best.clamp(min=0)
                g = torch.stack([best.to(torch.float64), self.d_bsn[idx], self.d_sde[idx],
                                 self.bw[bi].to(torch.float64), self.bl[bi].to(torch.float64),
                                 self.be[bi].to(torch.float64), self.h[bi]])
                hr = self.res_host[k].get(g.numel())
                hr.copy_(g.reshape(-1), non_blocking=True)
                if self.cfg.periodogram:
                    for s in plan.done:
                        soff, n, poff, m = plan.slot_off[s]
                        pg = {}
                        for name, arr in (("sn", self.sn), ("h", self.h), ("w", self.bw),
                                          ("l", self.bl), ("e", self.be)):
                            hp = torch.empty(m, dtype=arr.dtype, pin_memory=True)
                            hp.copy_(arr[poff:poff + m], non_blocking=True)
                            pg[name] = hp
                        spans.append(pg)
            ev = torch.cuda.Event()
            ev.record(st)
        self.nsteps += 1
        return (k, list(plan.done), ev, spans)

    def _run_graphs(self, trials: int):
        import torch
        parts = (self._rec_pre, self._rec_fold, self._rec_post)
        if self.cfg.use_graph and self.graphs is None:
            # the capture does not execute the kernels: this step runs eagerly
            torch.cuda.synchronize()
            gs = []
            for f in parts:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):  # captured on torch's side stream
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

    def collect(self, handle):
        """Wait for a launched step; returns {slot: (best index, S/N, SDE,
        w, l, e, H, periodogram or None)} of its finished requests."""
        k, done, ev, spans = handle
        ev.synchronize()
        nd = len(done)
        out = {}
        if nd:
            v = self.res_host[k].buf[:7 * nd].numpy().reshape(7, nd).copy()
            for j, s in enumerate(done):
                pg = None
                if spans:
                    pg = {key: t.numpy().copy() for key, t in spans[j].items()}
                out[s] = (int(v[0, j]), float(v[1, j]), float(v[2, j]), int(v[3, j]),
                          int(v[4, j]), int(v[5, j]), float(v[6, j]), pg)
        return out

    def fold_stats(self):
        """(launches, mean fold kernel ms, total trials, total fold seconds) of
        the logged steps; synchronises."""
        import torch
        torch.cuda.synchronize()
        ms = [a.elapsed_time(b) for a, b, _ in self.fold_log]
        tr = sum(k for _, _, k in self.fold_log)
        return len(ms), (float(np.mean(ms)) if ms else 0.0), tr, sum(ms) * 1e-3


class TransitServer:
    def __init__(self, cfg: ServerConfig | None = None, backend=None):
        self.cfg = cfg = cfg or ServerConfig()
        self.dev = backend if backend is not None else HipBackend(cfg)
        self.scap, self.pcap = self.dev.scap, self.dev.pcap
        self.salloc = FirstFit(self.scap)
        self.palloc = FirstFit(self.pcap)
        self.free_slots = list(range(cfg.max_curves - 1, -1, -1))
        self.slots: dict[int, _Slot] = {}
        self.order: list[int] = []   # resident slots, oldest first
        self.queue: list[Request] = []
        self.steps = 0
        self.inflight: list = []     # launched, not collected: (handle, slots used, done slots)
        self.dead: set = set()       # failed or cancelled slots awaiting release
        self.cancelled: set = set()
        self._out: list[Result] = []

    # -- compatibility accessors ---------------------------------------------
    @property
    def fold_log(self):
        return self.dev.fold_log

    @fold_log.setter
    def fold_log(self, v):
        self.dev.fold_log = v

    def fold_stats(self):
        return self.dev.fold_stats()

    # -- admission -----------------------------------------------------------
    def submit(self, req: Request) -> None:
        validate(req, self.cfg.max_len)
        if len(req.t) > self.scap or len(req.periods) > self.pcap:
            raise ValueError(f"request {req.rid} exceeds the HBM budget")
        self.queue.append(req)

    def cancel(self, rid: int) -> bool:
        """Remove a queued or resident request; returns whether it was found.
        Its memory is released once no step in flight uses it; no result is
        returned for it."""
        for i, r in enumerate(self.queue):
            if r.rid == rid:
                self.queue.pop(i)
                return True
        for s, sl in self.slots.items():
            if sl.req.rid == rid and s not in self.dead:
                self.cancelled.add(s)
                self._kill(s)
                return True
        return False

    def _in_flight(self, s: int) -> bool:
        return any(s in used for _, used, _ in self.inflight)

    def _kill(self, s: int) -> None:
        """Stop scheduling slot s; release it now or when its steps complete."""
        self.dead.add(s)
        if s in self.order:
            self.order.remove(s)
        if not self._in_flight(s):
            self._release(s)

    def _release(self, s: int) -> None:
        sl = self.slots.pop(s)
        if s in self.order:
            self.order.remove(s)
        self.dead.discard(s)
        self.cancelled.discard(s)
        self.salloc.release(sl.soff, len(sl.req.t))
        self.palloc.release(sl.poff, len(sl.req.periods))
        self.dev.release(s)
        self.free_slots.append(s)

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
            try:
                self.dev.upload(s, so, po, req)
            except Exception as exc:  # this request alone fails
                self._out.append(_failed(req, f"upload failed: {exc}", time.perf_counter()))
                self._kill(s)
                continue
            new.append(s)
        return new

    # -- one step ------------------------------------------------------------
    def _plan(self, new) -> StepPlan:
        cfg = self.cfg
        segs_c, segs_s = [], []
        for s in new:
            n = len(self.slots[s].req.t)
            k = (n + cfg.seg_len - 1) // cfg.seg_len
            segs_c.append(np.full(k, s, np.int32))
            segs_s.append(np.arange(0, k * cfg.seg_len, cfg.seg_len, dtype=np.int32))
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
        if ic:
            ic, ip, cost = np.concatenate(ic), np.concatenate(ip), np.concatenate(cost)
            o = np.argsort(-cost, kind="stable")  # longest items first
            ic, ip = ic[o], ip[o]
        else:
            ic, ip = np.zeros(0, np.int32), np.zeros(0, np.int64)
        slot_off = {s: (sl.soff, len(sl.req.t), sl.poff, len(sl.req.periods))
                    for s, sl in self.slots.items()}
        return StepPlan(new, np.concatenate(segs_c) if segs_c else np.zeros(0, np.int32),
                        np.concatenate(segs_s) if segs_s else np.zeros(0, np.int32), ic, ip, done,
                        slot_off, nep * len(cfg.widths) * cfg.nl)

    def _collect_oldest(self) -> list[Result]:
        """Wait for the oldest step in flight and retire its finished requests."""
        handle, used, done = self.inflight.pop(0)
        got = self.dev.collect(handle)
        tdone = time.perf_counter()
        out = []
        for s in done:
            sl = self.slots[s]
            if s not in self.dead:
                best, bsn, sde, w, l_, e, h, pg = got[s]
                j = best - sl.poff if best >= 0 else -1
                out.append(Result(sl.req.rid,
                                  float(sl.req.periods[j]) if j >= 0 else float("nan"), j, bsn,
                                  sde, w if j >= 0 else -1, l_ if j >= 0 else -1,
                                  e if j >= 0 else -1, h if j >= 0 else float("nan"),
                                  sl.req.arrival, tdone, periodogram=pg))
            self._release(s)
        for s in list(self.dead):
            if s in self.slots and not self._in_flight(s):
                self._release(s)
        return out

    def step(self, now: float | None = None) -> list[Result]:
        """Admit arrived requests, plan and launch one device step, then
        collect the step launched before it (pipelined: the device always has
        the next step queued while the host collects and plans).  Returns the
        results that became available; :meth:`flush` collects the rest."""
        now = time.perf_counter() if now is None else now
        new = self._admit(now)
        plan = self._plan(new)
        out = self._out
        self._out = []
        if not (len(plan.item_curve) or plan.done or plan.new):
            out.extend(self.flush())
            return out
        used = set(plan.item_curve.tolist()) | set(plan.done) | set(new)
        try:
            handle = self.dev.launch(plan)
        except Exception as exc:
            # the step did not run: every request with work in it fails alone
            tnow = time.perf_counter()
            for s in sorted(used):
                if s in self.slots and s not in self.dead:
                    out.append(_failed(self.slots[s].req, f"step failed: {exc}", tnow))
                    self._kill(s)
            return out
        self.steps += 1
        self.inflight.append((handle, used, list(plan.done)))
        if len(self.inflight) > 1:
            out.extend(self._collect_oldest())
        return out

    def flush(self) -> list[Result]:
        """Collect every step in flight."""
        out = self._out
        self._out = []
        while self.inflight:
            out.extend(self._collect_oldest())
        for s in list(self.dead):
            if s in self.slots:
                self._release(s)
        return out

    def busy(self) -> bool:
        return bool(self.queue or self.slots)

    def run(self, requests, t0: float | None = None, realtime: bool = True,
            progress=None, every: float = 5.0) -> list[Result]:
        """Serve a trace.  ``arrival`` of each request is seconds after t0; with
        realtime=False every request is available immediately.  ``progress``,
        if given, is called with (results so far, requests queued) at most
        every ``every`` seconds."""
        t0 = time.perf_counter() if t0 is None else t0
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.arrival = t0 + (r.arrival if realtime else 0.0)
            self.submit(r)
        results = []
        last = time.perf_counter()
        while self.busy():
            if progress is not None and time.perf_counter() - last > every:
                last = time.perf_counter()
                progress(len(results), len(self.queue))
            now = time.perf_counter()
            pending = any(sl.next_item < len(sl.req.periods) for s, sl in self.slots.items()
                          if s not in self.dead)
            if self.queue and not pending and self.queue[0].arrival > now:
                results.extend(self.flush())
                time.sleep(min(1e-3, max(0.0, self.queue[0].arrival - time.perf_counter())))
                continue
            results.extend(self.step(now))
        results.extend(self.flush())
        return results
