"""Detection quality of the periodic trapezoid search on the device.

A result counts as a recovery when its best period lies within 1% of the
injected period or of P/3, P/2, 2P or 3P.  tls.py's statistic (the default
output) divides H by the per-sample residual RMS, so it does not grow with
the number of transits and its periodogram maximum need not sit at an
injected period; the opt-in least-squares statistic (stat="dchi2") is the
detection statistic.  The figures are deterministic for the fixed seeds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_single_curve_workload_recovery(gpu):
    """The single-curve workload (4096 samples, trapezoid of depth 2e-3 at
    P = 3.7 d, 1000 periods 0.5-20 d): the least-squares statistic's best
    period is the injected one; tls.py's S/N peaks elsewhere."""
    from gtrap_amd import synth, tlsfold
    d = synth.tls_single_curve()
    curve = [(d["t"], d["flux"])]
    ls = tlsfold.tlsfold(curve, [d["periods"]], d["widths"], d["nl"], synth.CADENCE,
                         stat="dchi2")[0]
    sn = tlsfold.tlsfold(curve, [d["periods"]], d["widths"], d["nl"], synth.CADENCE)[0]
    p_ls = d["periods"][int(np.nanargmax(ls["sn"]))]
    p_sn = d["periods"][int(np.nanargmax(sn["sn"]))]
    print(f"least squares: {p_ls:.4f} d, tls.py S/N: {p_sn:.4f} d, injected {d['truth']['P']} d",
          flush=True)
    assert abs(p_ls - d["truth"]["P"]) <= 0.01 * d["truth"]["P"]
    assert not synth.recovered(p_sn, d["truth"])


def test_serving_trace_recovery_fraction(gpu):
    """A serving trace (Poisson arrivals, 2048-70000 samples, trapezoids of
    depth 1e-3 to 5e-3 in 30% of the requests) served with the least-squares
    statistic: the fraction of injected requests recovered, and none of them
    with tls.py's S/N.  The period grid (geomspace, 500-4000 periods) is not
    dense enough for every injection: for about four in ten of them no grid
    period keeps the transit aligned over the baseline."""
    import dataclasses
    from gtrap_amd import synth
    from gtrap_amd.serve import ServerConfig, TransitServer
    reqs, truth = synth.serving_trace(120, seed=2024)
    cfg = ServerConfig.tuned()
    out = {}
    for stat in ("snr", "dchi2"):
        srv = TransitServer(dataclasses.replace(cfg, stat=stat))
        res = srv.run(synth.serving_trace(120, seed=2024)[0], realtime=False)
        assert len(res) == len(reqs) and all(r.ok for r in res)
        inj = [r for r in res if truth[r.rid] is not None]
        out[stat] = (len(inj), sum(synth.recovered(r.best_period, truth[r.rid]) for r in inj))
    print("recovered (injected, recovered):", out, flush=True)
    ninj, rec = out["dchi2"]
    assert ninj >= 30
    assert rec / ninj >= 0.5  # 19 of 32 on this trace
    assert out["snr"][1] <= 0.1 * ninj
