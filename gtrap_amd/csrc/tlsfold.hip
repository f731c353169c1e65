// Periodic trapezoid least squares (tls_fold) for gfx950.
//
// Statistic: gtrap/tls.py:44-79 (get_chi2Hmax, divx=True, sigma=1) with
// S/N = -H/sqrt(res/(n-3)) (tls.py:100), evaluated on time folded at each
// trial period.  Definition (identical in the oracle):
//   u = fmod(t, P) (float64, exact), epoch e < ceil(P/dtep) at t0 = e*dtep,
//   d = u - t0, or (u - P) - t0 if d >= P/2, or (u + P) - t0 if d < -P/2;
//   L = m * (0.45*W/nl); regions A..E of tls.py:46-50 on d; mean over
//   [-W/2, W/2]; a trial needs n > 3 samples in [-W/2, W/2) (n - 3 degrees
//   of freedom) and W < P (a window as long as the period would hold two
//   images of a sample), else it is NaN; best trial per period = first
//   maximum in (W, L, epoch) order with NaN skipped (tlsfit, tls.py:82-110).
//
// MI355X design.  One 512-thread workgroup owns one (curve, period) work item
// at a time (persistent grid, atomic ticket, per-block global workspace):
//  1. fold + bucket sort: u of every valid sample is counted into ~n/2
//     phase buckets (LDS atomics), scanned, scattered to the workspace as
//     16-byte records, and each bucket is ordered by (u, flux bits) in LDS
//     chunks, giving the folded curve sorted by phase;
//  2. epoch tiles: the union of the windows [t0 - Wmax/2, t0 + Wmax/2] of up
//     to 96 epochs is a contiguous range of the (cyclically extended) sorted
//     curve; it is staged in LDS with float64 prefix sums of
//     (x', x'^2, q, q^2, x'q), q = phase relative to the tile's central
//     epoch, x' = x - xref, and a two-cells-per-sample phase index;
//  3. one thread per (epoch, W, L) trial: the four region boundaries come from
//     the phase index and a fixed-depth branch-free bisection with the exact
//     float64 offset test; region sums are prefix differences shifted to the
//     epoch's coordinates; trials are ranked by sign(S/N)*(S/N)^2 (one
//     division).  Windows larger than the LDS tile are swept in chunks with
//     the boundary sums accumulated per (epoch, W, boundary).
#include "tlsfold_kern.h"
#include <mutex>

namespace gtrap {
TfTiles g_tiles;
}  // namespace gtrap

// Workspace: a 256-byte header (work ticket) plus one slot per workgroup with
// the phase-sorted curve (16-byte records: float64 phase, float32 flux), the
// bucket offsets and the boundary rows of a wide pass (tf_block_bytes; every
// block's slot 256-byte aligned).
extern "C" int64_t gtrap_tlsfold_ws_bytes(int32_t nmax, int32_t nblocks) {
  if (nmax < 1) nmax = 1;
  return 256 + gtrap::tf_block_bytes(nmax) * (int64_t)(nblocks > 0 ? nblocks : 1);
}

static int64_t tf_slot_bytes(int32_t nmax) { return (gtrap_tlsfold_ws_bytes(nmax, 1) - 256); }

extern "C" int gtrap_tlsfold(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                             int32_t *best_e, const float *t, const float *x, const int64_t *off,
                             const int32_t *len, int32_t nmax, const double *periods,
                             const int64_t *poff, int32_t ncurves, int64_t nptot,
                             const double *widths, int32_t nw, int32_t nl, double dtep,
                             int32_t stat, void *ws, int64_t ws_bytes, int32_t nblocks,
                             void *stream) {
  GTRAP_REQUIRE(snr && best_w && best_l && best_e && t && x && off && len && periods && poff &&
                    widths && ws,
                "gtrap_tlsfold: null pointer");
  GTRAP_REQUIRE(ncurves >= 1 && nptot >= 0, "gtrap_tlsfold: bad counts");
  GTRAP_REQUIRE(nw >= 1 && nw <= gtrap::kTfMaxNw, "gtrap_tlsfold: nw must be 1..%d",
                gtrap::kTfMaxNw);
  GTRAP_REQUIRE(nl >= 1 && nl <= gtrap::kTfMaxNl, "gtrap_tlsfold: nl must be 1..%d",
                gtrap::kTfMaxNl);
  GTRAP_REQUIRE(dtep > 0.0, "gtrap_tlsfold: dtep must be positive");
  GTRAP_REQUIRE(stat == GTRAP_TLS_SNR || stat == GTRAP_TLS_DCHI2,
                "gtrap_tlsfold: stat must be GTRAP_TLS_SNR or GTRAP_TLS_DCHI2");
  GTRAP_REQUIRE(nmax >= 1, "gtrap_tlsfold: nmax must be >= 1");
  GTRAP_REQUIRE(nblocks >= 1, "gtrap_tlsfold: nblocks must be >= 1");
  GTRAP_REQUIRE(ws_bytes >= gtrap_tlsfold_ws_bytes(nmax, nblocks),
                "gtrap_tlsfold: workspace too small");
  if (nptot == 0) return GTRAP_OK;
  size_t dyn = 0;
  int te_run = 0;
  if (const int rc = gtrap::tf_launch_cfg<false, 0>(nw, nl, &dyn, &te_run)) return rc;
  gtrap::TfArgs A{t, x, off, len, periods, poff, nullptr, ncurves, nptot, widths, nw, nl, dtep,
                  snr, hbest, best_w, best_l, best_e, (unsigned char *)ws,
                  tf_slot_bytes(nmax), nmax, nullptr, nullptr, nullptr, nullptr, te_run, stat,
                  gtrap::g_tiles.max_tile, gtrap::g_tiles.cells, gtrap::g_tiles.spb, nullptr,
                  nullptr};
  const int64_t g = nptot < nblocks ? nptot : nblocks;
  hipStream_t st = gtrap::as_stream(stream);
  hipLaunchKernelGGL(gtrap::tf_zero_ticket, dim3(1), dim3(1), 0, st, (int *)ws);
  hipLaunchKernelGGL((gtrap::tlsfold_kernel<false, 0>), dim3((unsigned)g), dim3(gtrap::kTfThreads),
                     dyn, st, A);
  return gtrap::check_launch("tlsfold_kernel");
}

extern "C" int gtrap_tlsfold_set_tiles(int32_t max_tile_epochs, int32_t cells_per_sample,
                                       int32_t samples_per_bucket, int32_t table_run_epochs) {
  GTRAP_REQUIRE(max_tile_epochs >= 1 && max_tile_epochs <= gtrap::kTfMaxTile,
                "gtrap_tlsfold_set_tiles: max_tile_epochs must be 1..%d", gtrap::kTfMaxTile);
  GTRAP_REQUIRE(cells_per_sample == 1 || cells_per_sample == 2,
                "gtrap_tlsfold_set_tiles: cells_per_sample must be 1 or 2");
  GTRAP_REQUIRE(samples_per_bucket >= 1 && samples_per_bucket <= 8,
                "gtrap_tlsfold_set_tiles: samples_per_bucket must be 1..8");
  GTRAP_REQUIRE(table_run_epochs >= 0 && table_run_epochs <= gtrap::kTfMaxTe,
                "gtrap_tlsfold_set_tiles: table_run_epochs must be 0..%d", gtrap::kTfMaxTe);
  gtrap::g_tiles = gtrap::TfTiles{max_tile_epochs, cells_per_sample, samples_per_bucket,
                                  table_run_epochs};
  return GTRAP_OK;
}

// ---- the same search with no allocation: a static device workspace --------
//
// For callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): the work ticket and the
// per-workgroup slots live in one file-scope device array of
// kTfStaticBytes per GPU, held from the library's load on that GPU until the
// process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (resolved once per
// device); the array's name is a host-side stand-in, never passed to a kernel.
namespace gtrap {
constexpr int64_t kTfStaticBytes = 64ll << 20;
constexpr int kTfMaxDevices = 64;
}  // namespace gtrap

static __device__ __align__(256) unsigned char scratch[gtrap::kTfStaticBytes];

namespace gtrap {
static void *g_scratch_ptr[kTfMaxDevices];
static hipError_t g_scratch_err[kTfMaxDevices];
static std::once_flag g_scratch_once[kTfMaxDevices];

// Device address of this GPU's scratch array (hipGetSymbolAddress, once per
// device, with that device current).
static int tf_scratch_ptr(void **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kTfMaxDevices) {
    set_error("gtrap_tlsfold_static: no current device (or device index >= %d)",
              kTfMaxDevices);
    return GTRAP_ELAUNCH;
  }
  std::call_once(g_scratch_once[dev], [dev]() {
    g_scratch_err[dev] = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev]) {
    set_error("gtrap_tlsfold_static: hipGetSymbolAddress failed on device %d: %s", dev,
              hipGetErrorString(g_scratch_err[dev]));
    return GTRAP_ELAUNCH;
  }
  *out = g_scratch_ptr[dev];
  return GTRAP_OK;
}
}  // namespace gtrap

extern "C" int gtrap_tlsfold_static(double *snr, double *hbest, int32_t *best_w,
                                    int32_t *best_l, int32_t *best_e, const float *t,
                                    const float *x, const int64_t *off, const int32_t *len,
                                    int32_t nmax, const double *periods, const int64_t *poff,
                                    int32_t ncurves, int64_t nptot, const double *widths,
                                    int32_t nw, int32_t nl, double dtep, int32_t stat,
                                    int32_t nblocks, void *stream) {
  GTRAP_REQUIRE(nmax >= 1 && nblocks >= 1, "gtrap_tlsfold_static: bad nmax or nblocks");
  const int64_t slot = tf_slot_bytes(nmax);
  GTRAP_REQUIRE(256 + slot <= gtrap::kTfStaticBytes,
                "gtrap_tlsfold_static: nmax = %d needs a %lld-byte workgroup slot; the static "
                "workspace holds %lld bytes (use gtrap_tlsfold with a caller workspace)",
                nmax, (long long)slot, (long long)gtrap::kTfStaticBytes);
  void *ws = nullptr;
  if (const int rc = gtrap::tf_scratch_ptr(&ws)) return rc;
  const int64_t fit = (gtrap::kTfStaticBytes - 256) / slot;
  const int32_t nb = (int64_t)nblocks < fit ? nblocks : (int32_t)fit;
  return gtrap_tlsfold(snr, hbest, best_w, best_l, best_e, t, x, off, len, nmax, periods, poff,
                       ncurves, nptot, widths, nw, nl, dtep, stat, ws, gtrap::kTfStaticBytes,
                       nb, stream);
}

extern "C" int gtrap_tlsfold_items(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                                   int32_t *best_e, const float *t, const float *x,
                                   const int64_t *off, const int32_t *len, int32_t nmax,
                                   const double *periods, const int32_t *item_curve,
                                   const int64_t *item_pidx, const int32_t *nitems_dev,
                                   int32_t *counter, const double *widths, int32_t nw,
                                   int32_t nl, double dtep, int32_t stat, void *ws,
                                   int64_t ws_bytes, int32_t nblocks, void *stream) {
  GTRAP_REQUIRE(snr && best_w && best_l && best_e && t && x && off && len && periods &&
                    item_curve && item_pidx && nitems_dev && counter && widths && ws,
                "gtrap_tlsfold_items: null pointer");
  GTRAP_REQUIRE(nw >= 1 && nw <= gtrap::kTfMaxNw, "gtrap_tlsfold_items: nw must be 1..%d",
                gtrap::kTfMaxNw);
  GTRAP_REQUIRE(nl >= 1 && nl <= gtrap::kTfMaxNl, "gtrap_tlsfold_items: nl must be 1..%d",
                gtrap::kTfMaxNl);
  GTRAP_REQUIRE(dtep > 0.0 && nmax >= 1 && nblocks >= 1, "gtrap_tlsfold_items: bad scalars");
  GTRAP_REQUIRE(stat == GTRAP_TLS_SNR || stat == GTRAP_TLS_DCHI2,
                "gtrap_tlsfold_items: stat must be GTRAP_TLS_SNR or GTRAP_TLS_DCHI2");
  GTRAP_REQUIRE(ws_bytes >= gtrap_tlsfold_ws_bytes(nmax, nblocks),
                "gtrap_tlsfold_items: workspace too small");
  size_t dyn = 0;
  int te_run = 0;
  if (const int rc = gtrap::tf_launch_cfg<true, 0>(nw, nl, &dyn, &te_run)) return rc;
  gtrap::TfArgs A{t, x, off, len, periods, nullptr, nullptr, 0, 0, widths, nw, nl, dtep,
                  snr, hbest, best_w, best_l, best_e, (unsigned char *)ws,
                  tf_slot_bytes(nmax), nmax, item_curve, item_pidx, nitems_dev, counter, te_run,
                  stat, gtrap::g_tiles.max_tile, gtrap::g_tiles.cells, gtrap::g_tiles.spb, nullptr,
                  nullptr};
  hipLaunchKernelGGL((gtrap::tlsfold_kernel<true, 0>), dim3((unsigned)nblocks),
                     dim3(gtrap::kTfThreads),
                     dyn, gtrap::as_stream(stream), A);
  return gtrap::check_launch("tlsfold_kernel<items>");
}

namespace gtrap {

// Periodogram statistics of one curve's periods [poff[c], poff[c]+np[c]):
// best = first maximum of S/N over the values that are neither NaN nor -inf
// (+inf, a perfect fit, can win), mean and standard deviation over the finite
// S/N values, SDE = (best - mean)/std (the BLS statistic of
// gtrap/getstat.py:114-118 applied to the TLS periodogram).
__global__ __launch_bounds__(256) void tlsstats_kernel(const double *__restrict__ sn,
                                                       const int64_t *__restrict__ poff,
                                                       const int32_t *__restrict__ np_,
                                                       const int32_t *__restrict__ curves,
                                                       const int32_t *__restrict__ ncurves_dev,
                                                       int64_t *best_pidx, double *best_sn,
                                                       double *sde) {
  const int k = blockIdx.x;
  if (k >= *ncurves_dev) return;
  const int c = curves[k];
  const int64_t a = poff[c], m = np_[c];
  double bs = -INFINITY, s1 = 0.0, s2 = 0.0, cnt = 0.0;
  int64_t bi = -1;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
    const double v = sn[a + i];
    // a perfect fit (+inf) may win; -inf marks a period without any trial
    if (v > -INFINITY && (bi < 0 || v > bs)) { bs = v; bi = i; }
    if (isfinite(v)) {
      s1 += v;
      s2 += v * v;
      cnt += 1.0;
    }
  }
  __shared__ double r_bs[256], r_s1[256], r_s2[256], r_c[256];
  __shared__ int64_t r_bi[256];
  const int t = threadIdx.x;
  r_bs[t] = bs; r_bi[t] = bi; r_s1[t] = s1; r_s2[t] = s2; r_c[t] = cnt;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      const double ob = r_bs[t + s];
      const int64_t oi = r_bi[t + s];
      if (oi >= 0 && (r_bi[t] < 0 || ob > r_bs[t] || (ob == r_bs[t] && oi < r_bi[t]))) {
        r_bs[t] = ob;
        r_bi[t] = oi;
      }
      r_s1[t] += r_s1[t + s];
      r_s2[t] += r_s2[t + s];
      r_c[t] += r_c[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double mean = r_s1[0] / r_c[0];
    const double var = r_s2[0] / r_c[0] - mean * mean;
    best_pidx[c] = r_bi[0] >= 0 ? a + r_bi[0] : -1;
    best_sn[c] = r_bs[0];
    sde[c] = (r_bs[0] - mean) / sqrt(var);
  }
}

}  // namespace gtrap

extern "C" int gtrap_tlsstats(const double *sn, const int64_t *poff, const int32_t *np_,
                              const int32_t *curves, const int32_t *ncurves_dev,
                              int32_t max_curves, int64_t *best_pidx, double *best_sn,
                              double *sde, void *stream) {
  GTRAP_REQUIRE(sn && poff && np_ && curves && ncurves_dev && best_pidx && best_sn && sde,
                "gtrap_tlsstats: null pointer");
  GTRAP_REQUIRE(max_curves >= 0, "gtrap_tlsstats: bad max_curves");
  if (max_curves == 0) return GTRAP_OK;
  hipLaunchKernelGGL(gtrap::tlsstats_kernel, dim3((unsigned)max_curves), dim3(256), 0,
                     gtrap::as_stream(stream), sn, poff, np_, curves, ncurves_dev, best_pidx,
                     best_sn, sde);
  return gtrap::check_launch("tlsstats_kernel");
}
