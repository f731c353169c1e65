"""The no-allocation tls_fold entry point (static device workspace) against
the caller-workspace entry point, through the C ABI: identical outputs, on
two calls in a row, the second with a smaller input."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(lens, nper, seed):
    from gtrap_amd import tlsfold
    rng = np.random.default_rng(seed)
    curves, plist = [], []
    for n, m in zip(lens, nper):
        t = (np.arange(n) * 0.0204).astype(np.float32)
        t[rng.random(n) < 0.02] = -1000.0
        x = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
        curves.append((t, x))
        plist.append(np.geomspace(0.5, n * 0.0204 / 3, m))
    return tlsfold.pack(curves, plist)


def _outputs(b):
    import torch
    return [torch.full((b.nptot,), -7.0, dtype=torch.float64, device="cuda"),
            torch.full((b.nptot,), -7.0, dtype=torch.float64, device="cuda")] + \
        [torch.full((b.nptot,), -7, dtype=torch.int32, device="cuda") for _ in range(3)]


def _call_ws(b, widths, nblocks):
    import torch
    from gtrap_amd import _native
    p = _native.ptr
    need = int(_native.load().gtrap_tlsfold_ws_bytes(b.nmax, nblocks))
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    out = _outputs(b)
    _native.call("gtrap_tlsfold", *(p(o) for o in out), p(b.t), p(b.x), p(b.off), p(b.length),
                 b.nmax, p(b.periods), p(b.poff), b.ncurves, b.nptot, p(widths),
                 int(widths.numel()), 10, 0.0204, 0, p(ws), need, nblocks, _native.stream_handle())
    return out


def _call_static(b, widths, nblocks):
    from gtrap_amd import _native
    p = _native.ptr
    out = _outputs(b)
    _native.call("gtrap_tlsfold_static", *(p(o) for o in out), p(b.t), p(b.x), p(b.off),
                 p(b.length), b.nmax, p(b.periods), p(b.poff), b.ncurves, b.nptot, p(widths),
                 int(widths.numel()), 10, 0.0204, 0, nblocks, _native.stream_handle())
    return out


def test_static_workspace_matches_caller_workspace(gpu):
    import torch
    widths = torch.from_numpy(np.linspace(0.04, 0.3, 10)).cuda()
    # first call: long curves (the slot count clamps the 512 workgroups to 30),
    # second call: a smaller batch on the same static array
    for lens, nper, seed in (((70000, 9000, 2500), (40, 60, 30), 1), ((3000, 800), (25, 9), 2)):
        b = _batch(lens, nper, seed)
        print(f"no-alloc: curves {lens}, {b.nptot} periods", flush=True)
        want = _call_ws(b, widths, 512)
        got = _call_static(b, widths, 512)
        torch.cuda.synchronize()
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g.cpu().numpy(), w.cpu().numpy())
        assert np.all(got[2].cpu().numpy() >= -1)


def test_static_workspace_rejects_oversized_curves(gpu):
    import torch
    from gtrap_amd import _native
    widths = torch.from_numpy(np.linspace(0.04, 0.3, 10)).cuda()
    b = _batch((500,), (3,), 3)
    out = _outputs(b)
    p = _native.ptr
    with pytest.raises(ValueError, match="static"):
        _native.call("gtrap_tlsfold_static", *(p(o) for o in out), p(b.t), p(b.x), p(b.off),
                     p(b.length), 5_000_000, p(b.periods), p(b.poff), b.ncurves, b.nptot,
                     p(widths), 10, 10, 0.0204, 0, 64, _native.stream_handle())
