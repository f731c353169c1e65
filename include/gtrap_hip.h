/*
 * gtrap_hip.h -- C ABI of libgtrap_hip.so, the MI355X (gfx950) transit-search engine.
 *
 * Every entry point takes plain device pointers (allocated by the caller), plain
 * sizes and a hipStream_t (passed as void*; NULL = the legacy default stream),
 * launches asynchronously and returns a status code.  Nothing throws across the
 * ABI.  Outputs are written in place; no entry point allocates device memory
 * except gtrap_tlsfold_plan_bytes-sized workspaces the caller passes in.
 *
 * Image layout ("interleaved"): a batch of nq light curves of n samples each is
 * stored C-order as an (n, nq) float32 array, element (i, q) at i*nq + q.  This
 * is the layout the reference builds with np.array(lc).transpose()
 * (gtrap/read_keplerlc.py:49-50) and indexes in every kernel
 * (e.g. include/geebls.h:64,75).
 */
#ifndef GTRAP_HIP_H
#define GTRAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GTRAP_OK              0
#define GTRAP_EINVAL         -1  /* invalid argument (shape, size, constraint)  */
#define GTRAP_ELAUNCH        -2  /* kernel launch / HIP runtime error             */
#define GTRAP_ENOMEM         -3  /* workspace too small / allocation failure      */
#define GTRAP_EUNSUPPORTED   -4  /* valid request outside what this build serves  */

/* Human-readable text of the last error on the calling thread. */
const char *gtrap_last_error(void);

/* Library version string. */
const char *gtrap_version(void);

/* ---- median detrend (gfilter) ------------------------------------------ */
#define GTRAP_MEDFILT_SUB 0   /* out = x - median  (gtrap/gfilter.py:76)              */
#define GTRAP_MEDFILT_DIV 1   /* out = x / median  (Readme.md:50; IEEE fp32 divide)    */

/*
 * Replaces the gfilter kernel of gtrap/gfilter.py:13-85 (launched at
 * gtrap/gfilter.py:131).  For every row i < rows_out and column q < cols_out:
 *   med = median{ x~[i+d, q] : |d| <= r },  x~ = 0 outside 0 <= i < n
 *   out[i, q] = x[i, q] - med   (mode SUB)   or   x[i, q] / med   (mode DIV)
 * Rows/columns outside the covered range are left untouched, as the reference
 * leaves rows beyond grid.y*block.y unwritten.  Layout: interleaved (n, nq).
 * seg_len: samples per work item along time (execution parameter, 0 = auto).
 */
int gtrap_medfilt(const float *in, float *out, uint32_t n, uint32_t nq, uint32_t r,
                  int mode, uint32_t rows_out, uint32_t cols_out, uint32_t seg_len,
                  void *stream);

/*
 * Replaces the host-side offset of gtrap/gfilter.py:94-101 (get_offset):
 *   offset[q] = (sum_i [tu[i,q] >= 0] * img[i,q]) / n      (division by n, not ntrue)
 * accumulated in float64 on the device; also writes ntrue[q] = #{i : tu[i,q] >= 0}
 * when ntrue != NULL.  offset64 may be NULL; offset32 receives the float32 cast.
 */
int gtrap_offset(const float *img, const float *tu, uint32_t n, uint32_t nq,
                 double *offset64, float *offset32, uint32_t *ntrue, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GTRAP_HIP_H */
