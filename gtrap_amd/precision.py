"""Precision study of the BLS search: storage, accumulation and fold arithmetic.

Each variant runs the same box search (include/geebls.h:1-167) with one
arithmetic choice changed (``gtrap_geebls_prec``):

==========  ======================================================================
storage     ``f32`` (lc - offset in float32) or ``f16`` (binary16 after
            (lc - offset) * 2^k per curve)
accum       ``atomic32`` (float32 LDS atomics + float32 running sum, the
            reference's arithmetic), ``exact32`` (exact bin sums rounded to
            float32 + float32 running sum, gtrap_geebls), ``twosum32``
            (exact bins + compensated float32 running sum), ``f64``
fold        ``f32`` (frac(tu*f) in float32 from float32 time, the reference's)
            or ``f64`` (float64 time and frequency)
==========  ======================================================================

The float64 variant (f32 storage, f64 accumulation, f64 fold) is the yardstick:
SR per (curve, frequency) and SDE per curve of every other variant are
reported as errors against it.
"""
from __future__ import annotations

import numpy as np

from . import _native

ACC = {"atomic32": 0, "exact32": 1, "twosum32": 2, "f64": 3}
TRUTH = ("f32", "f64", "f64")

# (storage, accumulation, fold): one axis changed from the float64 yardstick,
# then the reference's arithmetic and gtrap_geebls's
STUDY = (
    ("f32", "atomic32", "f64"), ("f32", "exact32", "f64"), ("f32", "twosum32", "f64"),
    ("f32", "f64", "f32"), ("f16", "f64", "f64"),
    ("f32", "atomic32", "f32"), ("f32", "exact32", "f32"), ("f16", "exact32", "f64"),
)


def variant_code(storage: str, accum: str, fold: str) -> int:
    if storage not in ("f32", "f16") or accum not in ACC or fold not in ("f32", "f64"):
        raise ValueError(f"unknown variant {(storage, accum, fold)}")
    return (1 if storage == "f16" else 0) | (ACC[accum] << 1) | ((1 if fold == "f64" else 0) << 3)


def to_f16(lc: np.ndarray, offset: np.ndarray):
    """binary16 storage: (lc - offset) in float32, times 2^k per curve with
    max |.| * 2^k < 2^14; returns (uint16 bits (n, nq), 2^-k per curve)."""
    v = lc - offset[None, :].astype(np.float32)
    mx = np.abs(v).max(0)
    e = np.frexp(np.where(mx > 0, mx, 1.0))[1]
    k = 14 - e
    h = (v * np.ldexp(np.float32(1.0), k)[None, :].astype(np.float32)).astype(np.float16)
    return h.view(np.uint16), np.ldexp(1.0, -k).astype(np.float32)


class Inputs:
    """Device copies of a precision workload (synth.precision_batch layout)."""

    def __init__(self, d: dict):
        import torch
        _native.require_gpu()
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        lc = d["lc"]
        n, nq = lc.shape
        self.n, self.nq = n, nq
        t64 = np.asarray(d["t64"], np.float64)
        self.lc = dev(lc)
        self.tu = dev(np.broadcast_to(t64.astype(np.float32)[:, None], (n, nq)))
        self.t64 = dev(np.broadcast_to(t64[:, None], (n, nq)))
        self.offset = dev(np.asarray(d["offset"], np.float32))
        self.ntrue = dev(np.asarray(d["ntrue"], np.uint32).view(np.int32))
        h, hinv = to_f16(lc, np.asarray(d["offset"], np.float32))
        self.lch = dev(h.view(np.int16))
        self.hinv = dev(hinv)
        self.grid = d["grid"]


def run_variant(inp: Inputs, storage: str, accum: str, fold: str, stream=None):
    """SR (nq, nf) float64 device tensor and (start, end) bins of one variant."""
    import torch
    g = inp.grid
    nf = g["nf"]
    sr = torch.empty(nf * inp.nq, dtype=torch.float64, device="cuda")
    p1 = torch.empty(nf * inp.nq, dtype=torch.int32, device="cuda")
    p2 = torch.empty(nf * inp.nq, dtype=torch.int32, device="cuda")
    p = _native.ptr
    _native.call("gtrap_geebls_prec", variant_code(storage, accum, fold), p(sr), p(p1), p(p2),
                 p(inp.lc), p(inp.lch), p(inp.hinv), p(inp.tu), p(inp.t64), p(inp.offset),
                 p(inp.ntrue), inp.n, inp.nq, nf, g["nb"], g["kma"], g["kmi"], g["kkmi"],
                 float(g["fmin"]), float(g["df"]), g["nthread"], _native.stream_handle(stream))
    return sr.view(inp.nq, nf), p1.view(inp.nq, nf), p2.view(inp.nq, nf)


def sde(sr: np.ndarray) -> np.ndarray:
    """(max - mean) / std over each curve's frequencies, float64."""
    sr = np.asarray(sr, np.float64)
    return (sr.max(1) - sr.mean(1)) / sr.std(1)


def compare(sr: np.ndarray, ref: np.ndarray) -> dict:
    """SR and SDE errors of one variant against the yardstick."""
    rel = np.abs(sr - ref) / np.maximum(np.abs(ref), 1e-300)
    s, r = sde(sr), sde(ref)
    return dict(sr_rel_median=float(np.median(rel)), sr_rel_p99=float(np.percentile(rel, 99)),
                sr_rel_max=float(rel.max()), sr_abs_max=float(np.abs(sr - ref).max()),
                sde_abs_max=float(np.abs(s - r).max()),
                sde_rel_max=float((np.abs(s - r) / np.abs(r)).max()),
                best_freq_agree=float(np.mean(sr.argmax(1) == ref.argmax(1))))


def study(d: dict, variants=STUDY, timing: bool = True) -> dict:
    """Runs the yardstick and every variant; returns the error table and the
    kernel time of each variant (HIP events on the current stream)."""
    import torch
    inp = Inputs(d)
    out = {}

    def timed(v):
        run_variant(inp, *v)  # warm (code object load)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        sr, _, _ = run_variant(inp, *v)
        b.record()
        torch.cuda.synchronize()
        return sr.cpu().numpy(), a.elapsed_time(b)

    ref, tref = timed(TRUTH)
    out["f32/f64/f64 (yardstick)"] = dict(ms=tref, best_snr_sde_median=float(np.median(sde(ref))))
    for v in variants:
        sr, ms = timed(v)
        out["/".join(v)] = dict(ms=ms, **compare(sr, ref))
    return out


# ----------------------------------------------------------------------------
# the periodic trapezoid search (tls_fold) under the same three axes
# ----------------------------------------------------------------------------
TLS_VARIANTS = {  # name: variant bits of gtrap_tlsfold_prec
    "f32/f64/f64": 0,          # production: float32 flux, float64 sums and fold
    "f16/f64/f64": 1,          # binary16 storage after offset and scale
    "f32/f32/f64": 4,          # float32 prefix sums
    "f32/twosum32/f64": 8,     # float32 prefix sums with two-sum compensation
    "f32/f64/f32": 2,          # float32 fold
    "f16/twosum32/f64": 9,     # binary16 storage, compensated float32 sums, float64 fold
    "f32/f32/f32": 6,          # float32 throughout (the statistic aside)
}


class TlsInputs:
    """Device copies of a tls_fold precision workload: curves (t, flux) of one
    length, binary16 copies after removing each curve's mean and scaling by a
    power of two, and the period grids."""

    def __init__(self, curves, periods_list):
        import torch
        from .tlsfold import pack
        _native.require_gpu()
        self.batch = pack(curves, periods_list)
        x = np.concatenate([np.asarray(f, np.float32) for _, f in curves])
        hs, scales = [], []
        for _, f in curves:
            f = np.asarray(f, np.float32)[:, None]
            h, hinv = to_f16(f, np.array([f.mean(dtype=np.float64)], np.float32))
            hs.append(h[:, 0])
            scales.append(hinv[0])
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        self.x16 = dev(np.concatenate(hs).view(np.int16))
        self.x16scale = dev(np.array(scales, np.float32))
        assert self.x16.numel() == x.size


def tls_run_variant(inp: TlsInputs, variant: int, widths, nl=10, dtep=0.0204, stat="snr",
                    nblocks=512):
    """Per-period (sn, h, w, l, e) device tensors of one arithmetic variant."""
    import torch
    from .tlsfold import STATS, Workspace
    b = inp.batch
    nb = max(1, min(int(nblocks), b.nptot))
    ws = getattr(inp, "_ws", None) or Workspace()
    inp._ws = ws
    buf, need = ws.get(b.nmax, nb)
    wdev = torch.from_numpy(np.ascontiguousarray(widths, np.float64)).cuda()
    out = dict(sn=torch.empty(b.nptot, dtype=torch.float64, device="cuda"),
               h=torch.empty(b.nptot, dtype=torch.float64, device="cuda"),
               w=torch.empty(b.nptot, dtype=torch.int32, device="cuda"),
               l=torch.empty(b.nptot, dtype=torch.int32, device="cuda"),
               e=torch.empty(b.nptot, dtype=torch.int32, device="cuda"))
    p = _native.ptr
    _native.call("gtrap_tlsfold_prec", int(variant), p(out["sn"]), p(out["h"]), p(out["w"]),
                 p(out["l"]), p(out["e"]), p(b.t), p(b.x), p(inp.x16), p(inp.x16scale),
                 p(b.off), p(b.length), int(b.nmax), p(b.periods), p(b.poff), int(b.ncurves),
                 int(b.nptot), p(wdev), int(wdev.numel()), int(nl), float(dtep), STATS[stat],
                 p(buf), int(need), nb, _native.stream_handle())
    return out


def periodogram_stats(sn: np.ndarray):
    """(best index, best value, SDE) of one curve's per-period values, the
    statistics of gtrap_tlsstats: first maximum over values that are neither
    NaN nor -inf, mean and standard deviation over the finite values."""
    sn = np.asarray(sn, np.float64)
    cand = ~np.isnan(sn) & ~np.isneginf(sn)
    if not cand.any():
        return -1, -np.inf, np.nan
    v = np.where(cand, sn, -np.inf)
    b = int(np.argmax(v))
    fin = sn[np.isfinite(sn)]
    with np.errstate(all="ignore"):
        mean = fin.mean()
        std = np.sqrt((fin * fin).mean() - mean * mean)
    return b, float(v[b]), float((v[b] - mean) / std)


def tls_study(curves, periods_list, widths, nl=10, dtep=0.0204, stat="snr",
              variants=TLS_VARIANTS) -> dict:
    """Every variant against the float64 production kernel: relative error of
    the per-period value, per-curve SDE error, the fraction of periods whose
    best (W, L, epoch) and of curves whose best period are unchanged, and the
    kernel milliseconds (HIP events, after a warm-up launch)."""
    import torch
    inp = TlsInputs(curves, periods_list)
    poff = inp.batch.poff_host
    out, ref = {}, None
    for name, var in variants.items():
        tls_run_variant(inp, var, widths, nl, dtep, stat)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        r = tls_run_variant(inp, var, widths, nl, dtep, stat)
        b.record()
        torch.cuda.synchronize()
        r = {k: v.cpu().numpy() for k, v in r.items()}
        ms = a.elapsed_time(b)
        if ref is None:
            ref = r
            out[name + " (yardstick)"] = dict(ms=ms)
            continue
        fin = np.isfinite(ref["sn"]) & np.isfinite(r["sn"])
        rel = np.abs(r["sn"][fin] - ref["sn"][fin]) / np.abs(ref["sn"][fin])
        same = (r["w"] == ref["w"]) & (r["l"] == ref["l"]) & (r["e"] == ref["e"])
        bp, sd = [], []
        for c in range(len(curves)):
            s0, s1 = poff[c], poff[c + 1]
            b0, _, d0 = periodogram_stats(ref["sn"][s0:s1])
            b1, _, d1 = periodogram_stats(r["sn"][s0:s1])
            bp.append(b0 == b1)
            sd.append(abs(d1 - d0) / abs(d0))
        out[name] = dict(ms=ms, sn_rel_median=float(np.median(rel)),
                         sn_rel_p99=float(np.percentile(rel, 99)), sn_rel_max=float(rel.max()),
                         trial_agree=float(same.mean()), best_period_agree=float(np.mean(bp)),
                         sde_rel_max=float(np.max(sd)))
    return out
