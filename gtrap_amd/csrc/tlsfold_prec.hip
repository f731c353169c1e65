// Precision study of the periodic trapezoid search (tls_fold): the kernel of
// tlsfold.hip with one arithmetic choice changed per variant, for measuring
// the error of single-precision designs against the float64 production
// kernel:
//   storage      float32 flux, or binary16 after removing an offset and
//                scaling by a power of two per curve (kVarF16);
//   accumulation the tile prefix sums in float64, float32 (kVarAcc32) or
//                float32 with two-sum compensation (kVarAccTS);
//   fold         t mod P in float64 (exact), or in float32 (kVarFold32).
// The statistic itself stays float64 in every variant (gtrap/tls.py:44-100
// semantics as in tlsfold.hip).
#include "tlsfold_kern.h"

namespace gtrap {

template <int kVar>
static int tf_prec_launch(const TfArgs &A0, int64_t g, hipStream_t st) {
  size_t dyn = 0;
  int te_run = 0;
  if (const int rc = tf_launch_cfg<false, kVar>(A0.nw, A0.nl, &dyn, &te_run)) return rc;
  TfArgs A = A0;
  A.te_run = te_run;
  hipLaunchKernelGGL(tf_zero_ticket, dim3(1), dim3(1), 0, st, reinterpret_cast<int *>(A.ws));
  hipLaunchKernelGGL((tlsfold_kernel<false, kVar>), dim3((unsigned)g), dim3(kTfThreads), dyn, st,
                     A);
  return check_launch("tlsfold_kernel (precision variant)");
}

}  // namespace gtrap

extern "C" int gtrap_tlsfold(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                             int32_t *best_e, const float *t, const float *x, const int64_t *off,
                             const int32_t *len, int32_t nmax, const double *periods,
                             const int64_t *poff, int32_t ncurves, int64_t nptot,
                             const double *widths, int32_t nw, int32_t nl, double dtep,
                             int32_t stat, void *ws, int64_t ws_bytes, int32_t nblocks,
                             void *stream);

extern "C" int gtrap_tlsfold_prec(int32_t variant, double *snr, double *hbest, int32_t *best_w,
                                  int32_t *best_l, int32_t *best_e, const float *t,
                                  const float *x, const uint16_t *x16, const float *x16scale,
                                  const int64_t *off, const int32_t *len, int32_t nmax,
                                  const double *periods, const int64_t *poff, int32_t ncurves,
                                  int64_t nptot, const double *widths, int32_t nw, int32_t nl,
                                  double dtep, int32_t stat, void *ws, int64_t ws_bytes,
                                  int32_t nblocks, void *stream) {
  using namespace gtrap;
  if (variant == 0)
    return gtrap_tlsfold(snr, hbest, best_w, best_l, best_e, t, x, off, len, nmax, periods, poff,
                         ncurves, nptot, widths, nw, nl, dtep, stat, ws, ws_bytes, nblocks,
                         stream);
  GTRAP_REQUIRE(snr && best_w && best_l && best_e && t && off && len && periods && poff &&
                    widths && ws,
                "gtrap_tlsfold_prec: null pointer");
  GTRAP_REQUIRE(!(variant & kVarF16) || (x16 && x16scale), "gtrap_tlsfold_prec: binary16 flux "
                "and scales are required by this variant");
  GTRAP_REQUIRE((variant & kVarF16) || x, "gtrap_tlsfold_prec: null flux");
  GTRAP_REQUIRE(ncurves >= 1 && nptot >= 0 && nmax >= 1 && nblocks >= 1 && dtep > 0.0,
                "gtrap_tlsfold_prec: bad sizes");
  GTRAP_REQUIRE(nw >= 1 && nw <= kTfMaxNw && nl >= 1 && nl <= kTfMaxNl,
                "gtrap_tlsfold_prec: nw must be 1..%d and nl 1..%d", kTfMaxNw, kTfMaxNl);
  GTRAP_REQUIRE(stat == GTRAP_TLS_SNR || stat == GTRAP_TLS_DCHI2, "gtrap_tlsfold_prec: bad stat");
  GTRAP_REQUIRE(ws_bytes >= gtrap_tlsfold_ws_bytes(nmax, nblocks),
                "gtrap_tlsfold_prec: workspace too small");
  if (nptot == 0) return GTRAP_OK;
  const int64_t slot = gtrap_tlsfold_ws_bytes(nmax, 1) - 256;
  TfArgs A{t, x, off, len, periods, poff, nullptr, ncurves, nptot, widths, nw, nl, dtep,
           snr, hbest, best_w, best_l, best_e, (unsigned char *)ws, slot, nmax, nullptr,
           nullptr, nullptr, nullptr, 0, stat, g_tiles.max_tile, g_tiles.cells, g_tiles.spb,
           x16, x16scale};
  const int64_t g = nptot < nblocks ? nptot : nblocks;
  hipStream_t st = as_stream(stream);
  switch (variant) {
    case kVarF16: return tf_prec_launch<kVarF16>(A, g, st);
    case kVarAcc32: return tf_prec_launch<kVarAcc32>(A, g, st);
    case kVarAccTS: return tf_prec_launch<kVarAccTS>(A, g, st);
    case kVarFold32: return tf_prec_launch<kVarFold32>(A, g, st);
    case kVarF16 | kVarAccTS: return tf_prec_launch<kVarF16 | kVarAccTS>(A, g, st);
    case kVarAcc32 | kVarFold32: return tf_prec_launch<kVarAcc32 | kVarFold32>(A, g, st);
    default:
      set_error("gtrap_tlsfold_prec: unknown variant %d", variant);
      return GTRAP_EINVAL;
  }
}
