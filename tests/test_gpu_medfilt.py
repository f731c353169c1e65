"""GPU parity of the median detrend against the oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import medfilt as om

pytestmark = pytest.mark.gpu


def _run(img, r, mode, seg_len=0, rows_out=None, cols_out=None, fill=np.nan):
    import torch
    from gtrap_amd import gfilter
    n, nq = img.shape
    x = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    y = torch.full_like(x, fill)
    gfilter.medfilt_device(x, y, n, nq, r, mode, rows_out=rows_out, cols_out=cols_out,
                           seg_len=seg_len)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("r", [0, 1, 4, 7, 16, 24, 33, 48, 64, 65, 100])
@pytest.mark.parametrize("mode", ["sub", "div"])
def test_parity_random(gpu, r, mode):
    rng = np.random.default_rng(100 + r)
    img = (1.0 + 1e-3 * rng.standard_normal((1000, 70))).astype(np.float32)
    img[rng.random(img.shape) < 0.01] += 1e-2
    got = _run(img, r, mode)
    want = om.detrend_scipy(img, r, mode)
    np.testing.assert_array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("shape", [(513, 1), (64, 3), (5, 2), (1, 1), (300, 130)])
@pytest.mark.parametrize("r", [3, 24, 64, 70])
def test_parity_shapes_and_edges(gpu, shape, r):
    rng = np.random.default_rng(sum(shape) + r)
    img = rng.integers(-4, 5, size=shape).astype(np.float32)  # heavy ties
    for mode in ("sub", "div"):
        got = _run(img, r, mode)
        want = om.detrend_scipy(img, r, mode)
        np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=f"{shape} {r} {mode}")


@pytest.mark.parametrize("seg_len", [1, 7, 100, 5000])
@pytest.mark.parametrize("r", [16, 64, 100])
def test_segmentation_invariance(gpu, seg_len, r):
    """One-lane windows (r = 16), the blocked selection (r = 64: segments that
    are and are not whole 32-row blocks, blocks at the curve's ends) and
    windows shared by a lane group (r = 100)."""
    rng = np.random.default_rng(3)
    img = (1.0 + 1e-3 * rng.standard_normal((777, 9))).astype(np.float32)
    got = _run(img, r, "sub", seg_len=seg_len)
    want = om.detrend_scipy(img, r, "sub")
    np.testing.assert_array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("r", [10, 64])
def test_partial_coverage_leaves_rest_untouched(gpu, r):
    rng = np.random.default_rng(4)
    img = (1.0 + 1e-3 * rng.standard_normal((1000, 8))).astype(np.float32)
    got = _run(img, r, "sub", rows_out=900, cols_out=6, fill=-7.0)
    want = om.detrend_scipy(img, r, "sub")
    np.testing.assert_array_equal(_bits(got[:900, :6]), _bits(want[:900, :6]))
    assert np.all(got[900:] == -7.0) and np.all(got[:, 6:] == -7.0)


def test_dropin_get_detrend_and_module(gpu):
    """Reference call shapes: get_detrend(img, nby, nbx, r, osw) and the raw kernel."""
    from gtrap_amd import driver as cuda
    from gtrap_amd import gfilter
    rng = np.random.default_rng(5)
    n, nq, r = 2000, 4, 16
    img = np.asfortranarray((1 + 1e-3 * rng.standard_normal((n, nq))).astype(np.float32))
    host = gfilter.get_detrend(img, nby=1000, nbx=1, r=r, isw=0, osw=0)
    np.testing.assert_array_equal(_bits(host), _bits(om.detrend_scipy(img, r)))
    dev = gfilter.get_detrend(img, nby=1000, nbx=1, r=r, osw=1)
    out = np.zeros(n * nq, np.float32)
    cuda.memcpy_dtoh(out, dev)
    np.testing.assert_array_equal(_bits(out.reshape(n, nq)), _bits(host))
    # raw kernel launch exactly as gtrap/gfilter.py:131
    dev_in = cuda.mem_alloc(img.nbytes)
    cuda.memcpy_htod(dev_in, img.flatten().astype(np.float32))
    dev_out = cuda.mem_alloc(img.nbytes)
    gfilter.gmf_module().get_function("gfilter")(
        dev_in, dev_out, np.uint32(n), np.uint32(nq), np.uint32(r),
        block=(1, 1000, 1), grid=(nq, n // 1000), shared=4 * (2 * r + 1000))
    cuda.memcpy_dtoh(out, dev_out)
    np.testing.assert_array_equal(_bits(out.reshape(n, nq)), _bits(host))


def test_get_offset(gpu):
    from gtrap_amd import driver as cuda
    from gtrap_amd import gfilter
    rng = np.random.default_rng(6)
    n, nq = 3000, 5
    img = (1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    tu = np.tile(np.arange(n, dtype=np.float32)[:, None] * 0.02, (1, nq))
    tu[rng.random(tu.shape) < 0.05] = -1000.0
    dev = cuda.mem_alloc(img.nbytes)
    cuda.memcpy_htod(dev, img)
    offset, imgout = gfilter.get_offset(n, nq, dev, tu)
    ref = img.astype(np.float64).copy()
    ref[tu < 0] = 0.0
    np.testing.assert_allclose(offset, ref.mean(axis=0), rtol=1e-6, atol=1e-12)
    want_img = img.copy()
    want_img[tu < 0] = 0
    np.testing.assert_array_equal(imgout, want_img)
    off32, off64, ntrue = gfilter.get_offset(n, nq, dev, tu, device_only=True)
    np.testing.assert_allclose(off64.cpu().numpy(), ref.sum(0) / n, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(ntrue.cpu().numpy(), (tu >= 0).sum(0))


def test_detrend_image_full_size_subset(gpu):
    """4096 curves x 65536 samples, r in {7, 24, 64}: 16 random curves vs scipy,
    div mode (the workload's output is flux / median)."""
    import torch
    from gtrap_amd import gfilter, synth
    n, nq = 65536, 4096
    img = synth.detrend_image(nq=nq, n=n, device="cuda")
    out = torch.empty_like(img)
    rng = np.random.default_rng(11)
    cols = rng.choice(nq, 16, replace=False)
    sub = img[:, cols].cpu().numpy()
    for r in (7, 24, 64):
        gfilter.medfilt_device(img, out, n, nq, r, "div")
        torch.cuda.synchronize()
        got = out[:, cols].cpu().numpy()
        want = om.detrend_scipy(sub, r, "div")
        np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=f"r={r}")


@pytest.mark.parametrize("r", [65, 100, 128, 129, 200, 512, 700, 1031, 1032, 1500])
def test_wide_windows_of_the_reference_drivers(gpu, r):
    """The r of kepler/gtls_mockkepler.py:229 (128), kepler/gtls_pickkepler.py:197
    (100) and works/sp/gtls_tesslc.py:129 (512), and the edges of every lane
    group size (2, 4, 8, 16 lanes up to r = 128, 257, 515, 1031; the LDS
    bisection above): bit-exact against scipy.signal.medfilt, sub and div, with
    outliers and edges."""
    import torch
    from gtrap_amd import gfilter
    rng = np.random.default_rng(r)
    n, nq = 3000, 5
    img = (1.0 + 1e-3 * rng.standard_normal((n, nq))).astype(np.float32)
    img[rng.random((n, nq)) < 0.01] += np.float32(1e-2)
    x = torch.from_numpy(img).cuda()
    for mode in ("sub", "div"):
        y = torch.empty_like(x)
        gfilter.medfilt_device(x, y, n, nq, r, mode)
        want = om.detrend_scipy(img, r, mode)
        np.testing.assert_array_equal(y.cpu().numpy().view(np.uint32), want.view(np.uint32))
