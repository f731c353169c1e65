/*
 * gtrap_hip.h -- C ABI of libgtrap_hip.so, the MI355X (gfx950) transit-search engine.
 *
 * Every entry point takes plain device pointers (allocated by the caller), plain
 * sizes and a hipStream_t (passed as void*; NULL = the legacy default stream),
 * launches asynchronously and returns a status code.  Nothing throws across the
 * ABI.  Outputs are written in place; no entry point allocates device memory:
 * scratch space (gtrap_tlsfold) is a workspace the caller passes in.
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
#define GTRAP_EUNSUPPORTED   -4  /* valid request outside what the library serves */

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
 * Ragged-batch form of gtrap_medfilt for the serving loop (no reference
 * counterpart: the reference detrends one rectangular image per call).  Curve
 * c holds len[c] contiguous samples at x + off[c]; segment g filters rows
 * [seg_start[g], seg_start[g] + seg_len) of curve seg_curve[g] into y + off[c].
 * The segment count is read from *nseg_dev (<= max_segs) so the launch shape
 * is fixed and the call can be captured in a HIP graph.  r <= 4096 (r <= 64:
 * sorted windows in registers; larger r: LDS tiles of 256 rows, the r the
 * reference's drivers use: 100, 128, 512).
 */
int gtrap_medfilt_ragged(const float *x, float *y, const int64_t *off, const int32_t *len,
                         const int32_t *seg_curve, const int32_t *seg_start,
                         const int32_t *nseg_dev, int32_t max_segs, uint32_t r, int mode,
                         int32_t seg_len, void *stream);

/*
 * Replaces the host-side offset of gtrap/gfilter.py:94-101 (get_offset):
 *   offset[q] = (sum_i [tu[i,q] >= 0] * img[i,q]) / n      (division by n, not ntrue)
 * accumulated in float64 on the device; also writes ntrue[q] = #{i : tu[i,q] >= 0}
 * when ntrue != NULL.  offset64 may be NULL; offset32 receives the float32 cast.
 */
int gtrap_offset(const float *img, const float *tu, uint32_t n, uint32_t nq,
                 double *offset64, float *offset32, uint32_t *ntrue, void *stream);

/* ---- box least squares (geebls / geebls_simple) ------------------------- */

/*
 * Replaces the geebls kernel of include/geebls.h:1-167 (module factory
 * gtrap/geebls.py:10-22, MINBIN 5, INVT 0.0).  One trial frequency
 * f0 = fmin + df*jf per (jf < nf) and curve q < nq:
 *   fold samples i < floor(n/nthread)*nthread with tu >= 0 into nb phase bins,
 *   bin sums of (lc - offset[q]) and counts, wrap kma bins, scan start bin x width,
 *   power = s^2 / (kk (ntrue[q] - kk)) for width > max(kmi,1) and kk > max(kkmi,5),
 *   blsp = sqrt(max power), phase1/phase2 = start/end bin of the best box.
 * Outputs are curve-major: index jf + nf*q.  nthread is the reference's
 * blockDim.x (it decides the dropped sample tail, include/geebls.h:57).
 * phase1/phase2 may be NULL.  When no box qualifies, blsp = 0 and the phases are
 * 0xFFFFFFFF.  Ties resolve to the lowest (start, end) pair.
 */
int gtrap_geebls(float *blsp, uint32_t *phase1, uint32_t *phase2,
                 const float *lc, const float *tu, const float *offset,
                 const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf,
                 uint32_t nb, uint32_t kma, uint32_t kmi, uint32_t kkmi,
                 float fmin, float df, uint32_t nthread, void *stream);

/*
 * Replaces the geebls kernel of gtrap/geebls_simple.py:17-175: common 1-D time
 * array tu[n], no mask, no offset, N = n, blsp only.
 */
int gtrap_geebls_simple(float *blsp, const float *lc, const float *tu,
                        uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb,
                        uint32_t kma, uint32_t kmi, uint32_t kkmi,
                        float fmin, float df, uint32_t nthread, void *stream);

/*
 * gtrap_geebls and gtrap_geebls_simple with a caller workspace: no
 * counterpart in the reference (whose kernel reads the interleaved image in
 * place, include/geebls.h:60-77); for callers that can lend device memory.
 * One pass copies the interleaved (n, nq) image to curve-major order in the
 * workspace and takes each curve's fixed-point scale once, so the binning
 * loads are coalesced.  Results are bit-identical to the entry points above.
 * ws: at least gtrap_geebls_ws_bytes(n, nq, simple) bytes, 256-byte aligned;
 * GTRAP_ENOMEM when smaller.  Calls that share a workspace must not overlap.
 */
int64_t gtrap_geebls_ws_bytes(uint32_t n, uint32_t nq, int32_t simple);
int gtrap_geebls_ws(float *blsp, uint32_t *phase1, uint32_t *phase2,
                    const float *lc, const float *tu, const float *offset,
                    const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf,
                    uint32_t nb, uint32_t kma, uint32_t kmi, uint32_t kkmi,
                    float fmin, float df, uint32_t nthread, void *ws, int64_t ws_bytes,
                    void *stream);
int gtrap_geebls_simple_ws(float *blsp, const float *lc, const float *tu,
                           uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb,
                           uint32_t kma, uint32_t kmi, uint32_t kkmi,
                           float fmin, float df, uint32_t nthread, void *ws,
                           int64_t ws_bytes, void *stream);

/*
 * Arithmetic variants of the geebls kernel (include/geebls.h:1-167) for
 * precision studies; the search itself is unchanged.  variant bits:
 *   bit 0      flux read as binary16 lch[(n, nq)] = (lc - offset) * 2^k per
 *              curve, hinv[q] = 2^-k (else float32 lc - offset[q]);
 *   bits 1-2   accumulation: 0 float32 LDS atomics + float32 running sum
 *              (the reference's arithmetic), 1 exact bin sums rounded to
 *              float32 + float32 running sum (gtrap_geebls), 2 exact bins +
 *              two-sum float32 running sum, 3 exact bins + float64 running sum
 *              and statistic;
 *   bit 3      fold in float64 from t64[(n, nq)]: f = fmin + df*jf,
 *              ph = frac(t*f) (else the reference's float32 fold of tu).
 * Valid samples are tu >= 0 in every variant.  sr[jf + nf*q] in float64;
 * phase1/phase2 may be null.  Variant 2 (float32, exact, float32 fold) equals
 * gtrap_geebls bit for bit.
 */
int gtrap_geebls_prec(int variant, double *sr, uint32_t *phase1, uint32_t *phase2,
                      const float *lc, const uint16_t *lch, const float *hinv,
                      const float *tu, const double *t64, const float *offset,
                      const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf,
                      uint32_t nb, uint32_t kma, uint32_t kmi, uint32_t kkmi,
                      double fmin, double df, uint32_t nthread, void *stream);

/* ---- post-search peak statistics ----------------------------------------- */

/*
 * Replaces the statistics block after the search in gtrap/gtls.py:244-275 and
 * kepler/gtls_mockkepler.py:336-376, with the peak finder
 * gtrap/detect_peaks.py:10-138 (edge 'rising', minimum distance mpd), for
 * each series q of the interleaved sn[k*nq + q], k < nt (t0 alike, may be
 * null).  Per q: st[8q..8q+7] = maxsn, median, std (population, over
 * sn > 0), far, diff, pssn, pinterval, 0; cnt[2q] = peaks, cnt[2q+1] =
 * count(sn > 0); top[ntop*q + t] = the t-th highest peak's index (-1 past
 * the last).  Ties in height: the larger index first.  Workspace:
 * gtrap_peakstats_ws_bytes(nt, nblocks) bytes; nt <= 131072.
 */
int64_t gtrap_peakstats_ws_bytes(int32_t nt, int32_t nblocks);
int gtrap_peakstats(double *st, int32_t *cnt, int32_t *top, const float *sn, const float *t0,
                    int32_t nt, int32_t nq, int32_t mpd, int32_t ntop, void *ws,
                    int64_t ws_bytes, int32_t nblocks, void *stream);

/* ---- candidate cut-outs ----------------------------------------------------- */

/*
 * Replaces gtrap/picktrap.py:27-139 (pick_Wnormalized_cleaned_lc_direct) for
 * a batch of candidates c: curve cq[c], epoch ct0[c], width cw[c] on the
 * interleaved (n, nq) lc/tu image.  Output per candidate: lcsx[c*nx + i],
 * igapx[c*nx + i] (float64), prec[c] (1/0) and status[c] (0 ok, 1 no valid
 * time >= T0, 2 gap cleaning stuck, 3 window shorter than 2 slots, 4 window
 * longer than 2048 slots, 5 curve index outside [0, nq); not ok -> all ones,
 * as the reference's fallback).
 */
int gtrap_cutout(double *lcsx, double *igapx, int32_t *prec, int32_t *status, const float *lc,
                 const float *tu, int32_t n, int32_t nq, const int32_t *cq, const double *ct0,
                 const double *cw, int32_t ncand, double alpha, int32_t nx, double daytopix,
                 int32_t contcrit, void *stream);

/* ---- image assembly ---------------------------------------------------- */

/*
 * Replaces the array assembly of gtrap/read_keplerlc.py:23-57 (load_keplc)
 * and :123-146 (throw_kepintarray) for curves already in device memory:
 * curve q's samples are t/flux/cno[off[q] + i], i < len[q] (maxlen >= every
 * len[q]).  Output: interleaved (n, nq) float32 lc and tu; slot
 * j = cno[i] - cno[off[q]] gets (float)flux and (float)(t - t0) unless flux is
 * NaN (t0 = t[off[q]] if offt, else 0); other slots hold fillv / inval; gap
 * slots (tu < 0) then get the curve's minimum flux slot at even j and its
 * maximum at odd j.  ntrue[q] = slots with tu >= 0, hfill/lfill = those
 * extrema, t0[q].  Samples whose slot is outside [0, n) are skipped and
 * counted in the device int *nbad (the reference stops for j >= n).
 */
int gtrap_image_build(float *lc, float *tu, uint32_t *ntrue, float *hfill, float *lfill,
                      double *t0, const double *t, const double *flux, const int64_t *cno,
                      const int64_t *off, const int32_t *len, int32_t nq, int32_t n,
                      int32_t maxlen, float fillv, float inval, int offt, int32_t *nbad,
                      void *stream);

/* ---- synthetic injection ------------------------------------------------- */

/*
 * Replaces the gammagen kernel of gtrap/ggenmock.py:18-79: x[id] ~ Gamma(a,
 * rate b) for id < count (Marsaglia-Tsang, u^(1/a) boost for a <= 1, at most
 * 10000 proposals; a slot whose proposals all fail is NaN and counted in the
 * device int *nfail).  Random words: Philox4x32-10 keyed by seed, counter
 * (id, proposal), so x depends only on (seed, id).
 */
int gtrap_gamma(float *x, int64_t count, double a, double b, uint64_t seed, int32_t *nfail,
                void *stream);

/*
 * Mock transit injection (the role of gtrap/genmock.py:9-43 for trapezoids):
 * on the interleaved (n, nq) image, subtract (mode 0) or multiply by one
 * minus (mode 1) the dip of the gtrap/tls.py:2-15 trapezoid, per curve q:
 * depth[q] on |d| < W/4 - L/2, linear to 0 at |d| = W/4 + L/2 (W = width[q],
 * L = ell[q]), d the offset of tu from the nearest t0[q] + k period[q]
 * (period <= 0: one event).  Samples with tu < 0 are left unchanged.
 */
int gtrap_inject_trapezoid(float *lc, const float *tu, int32_t n, int32_t nq, const double *t0,
                           const double *period, const double *depth, const double *width,
                           const double *ell, int mode, void *stream);

/* ---- BLS statistics (getstat) -------------------------------------------- */

/*
 * Replaces gtrap/getstat.py:14-123 (getmax kernel + patch_stat + combine_stat):
 * per curve q, over the first npatch*1024 frequencies (npatch = floor(nf/1024)),
 *   maxv = max blsp, idx = lowest index of the max,
 *   Pest = 1/(idx*df + fmin), ave = sum/nf, var = sumsq/nf - ave^2,
 *   sde = (maxv - ave)/sqrt(var), phasebest = phase1/2[idx].
 * Sums in float64.  phase1/phase2/phasebest1/phasebest2 may be NULL.
 * patch: the reference's nthread (1024 by default, gtrap/getstat.py:64).
 */
int gtrap_blsstats(const float *blsp, const uint32_t *phase1, const uint32_t *phase2,
                   uint32_t nf, uint32_t nq, uint32_t patch, double df, double fmin,
                   double *pest, double *sde, uint32_t *phasebest1,
                   uint32_t *phasebest2, float *maxval, void *stream);

/* ---- single-transit-epoch trapezoid least squares (gtls) ------------------ */
#define GTRAP_GTLS_NORMAL   0  /* include/gtls_simple.h   */
#define GTRAP_GTLS_ABSOLUTE 1  /* include/gtls_absolute.h */
#define GTRAP_GTLS_PARITY   2  /* or-ed into mode: the float32 sequential scans */

/*
 * Replaces the gtls kernel of include/gtls_simple.h:1-406 and
 * include/gtls_absolute.h (module factory gtrap/gtls.py:11-27, FILLVAL -1000,
 * FILFACCRIT 0.75).  Epoch k < nt, curve q < nq, nw widths W = wmax - dw*iw
 * (the reference's blockDim.x, a power of two), nl ingress steps, scoop nsc.
 * Outputs (interleaved, index k*nq + q): tlssn, tlsw, tlst0, tlsl, tlshmax.
 * t0min and dt are accepted and unused, as in the reference.
 *
 * Two arithmetics, the same region decisions (shifted time fl32(T - deltat*k)
 * against the float32 boundaries, the region-A flux index quirk, the float32
 * fill factor):
 *   default: float64 sums, H, residual and S/N on the exact shifted time,
 *     O(log nsc) per (epoch, W, L) trial from tile prefix sums (nsc up to about
 *     2700); S/N is rounded to float32 at the end;
 *   mode | GTRAP_GTLS_PARITY: the reference's float32 sequential scans in its
 *     operation order, bit-identical to a float32 restatement (nsc up to
 *     19200).
 * The float32 arithmetic itself differs from the float64 one by a median of
 * a few 1e-6 and up to about 1e-2 relative in S/N (cancellation of
 * x - mean and of denA in float32).
 */
int gtrap_gtls(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
               const float *lc, const float *tu, int32_t nt, int32_t nq, int32_t nw,
               int32_t nl, int32_t nsc, float t0min, float dt, float wmax, float dw,
               float deltat, int mode, void *stream);

/* ---- trapezoid least squares over an explicit grid (tlsfit) -------------- */

/*
 * Replaces the triple loop of gtrap/tls.py:82-110 (tlsfit) over
 * get_chi2Hmax (tls.py:44-79): for every trial (w, l, k) (k fastest) with
 * W = Wl[w], L = Ll[l], t0 = tl[k] on the float64 series (t, x) of n samples:
 * H[trial] = Hmax, res[trial] = chi^2 / sigma^2, nout[trial] = samples in
 * [-W/2, W/2), xsum[trial] = mean over [-W/2, W/2] (0 when divx == 0).  NaN
 * samples are skipped in sums as numpy's nansum/nanmean.  xsum may be NULL.
 * sorted != 0 states that t is non-decreasing (no NaN): each trial then
 * visits only the samples of its window (same results, far less work).
 */
int gtrap_tlsgrid(const double *t, const double *x, int32_t n, const double *Wl, int32_t nW,
                  const double *Ll, int32_t nL, const double *tl, int32_t nT, int divx,
                  double sigma, int32_t sorted, double *H, double *res, int32_t *nout,
                  double *xsum, void *stream);

/*
 * tlsfit's choice over a gtrap_tlsgrid grid (replaces gtrap/tls.py:100-110:
 * sn = -H/sqrt(res/(n-3)), then np.nanargmax): sn per trial in float64 as
 * numpy evaluates it (written to sn when it is not NULL), and best[0..3] =
 * {index of the first maximum with NaN counted as -inf, its sn, its H, its
 * xsum (0 when xsum is NULL)}; best[0] = -1 when every sn is NaN (numpy
 * raises there).  All pointers are device memory of ntr trials, best of 4
 * doubles; one workgroup, asynchronous on stream.
 */
int gtrap_tlsfit_best(const double *H, const double *res, const int32_t *nout,
                      const double *xsum, int32_t ntr, double *sn, double *best, void *stream);

/* ---- periodic trapezoid least squares (tls_fold) --------------------------- */
#define GTRAP_TLS_SNR   0  /* tls.py's S/N = -H / sqrt(res/(n-3)) (tls.py:100) */
#define GTRAP_TLS_DCHI2 1  /* least-squares detection statistic, see gtrap_tlsfold */

/*
 * Periodic extension of the trapezoid statistic of gtrap/tls.py:44-110
 * (get_chi2Hmax with divx=True, tlsfit's S/N = -H/sqrt(res/(n-3)) and
 * nanargmax over the grid).  For every trial period P (float64, days) of a
 * curve: u = fmod(t, P), epochs e < ceil(P/dtep) at t0 = e*dtep, offset
 * d = u - t0 wrapped by +-P to [-P/2, P/2); widths W = widths[w] and
 * L = m*(0.45*W/nl), m < nl (the L grid of include/gtls_simple.h:69,167).
 * Samples with t < 0 (masked) or a non-finite flux are skipped.  A trial
 * with n <= 3 samples in its window or with W >= P is NaN.  Best trial per
 * period = first maximum in (w, m, e) order, NaN skipped.
 *
 * Ragged batch: curve c has len[c] samples at offset off[c] of the flat t, x
 * arrays; its periods are periods[poff[c] .. poff[c+1]) with poff[ncurves] =
 * nptot.  Outputs per period (index poff[c] + p): best S/N, its H, and the
 * w, m, e indices (-1 when every trial is NaN).
 * stat selects the ranking statistic, reported in snr: GTRAP_TLS_SNR (the
 * reference's S/N, default) or GTRAP_TLS_DCHI2, no counterpart in the
 * reference: the least-squares fit of the same trapezoid shape g (tls.py:2-15
 * with unit depth) to y = x - mean over the window [-W/2, W/2),
 * H* = sum(y g)/sum(g^2), D = sign(-H*) sqrt(dchi2 (n-3)/res*) with
 * dchi2 = sum(y g)^2/sum(g^2) and res* = sum(y^2) - dchi2 (a matched
 * filter); hbest then holds H*.  tls.py's S/N is H over the per-sample
 * residual RMS, so it does not grow with the number of transits folded into
 * a window; D does (about depth/sigma times the square root of the
 * in-transit samples), which makes it the statistic that recovers injected
 * periods.  L = 0 (m = 0) is no trial in both.
 * Precondition len[c] <= nmax: each workgroup's workspace slot holds nmax
 * samples.  The lengths live on the device, so the kernel checks it: the
 * periods of a longer curve are not searched and get S/N NaN, H NaN and
 * indices -2.
 * ws: caller workspace of at least gtrap_tlsfold_ws_bytes(nmax, nblocks)
 * bytes; nblocks = persistent workgroups (execution parameter).
 */
int64_t gtrap_tlsfold_ws_bytes(int32_t nmax, int32_t nblocks);
int gtrap_tlsfold(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                  int32_t *best_e, const float *t, const float *x, const int64_t *off,
                  const int32_t *len, int32_t nmax, const double *periods,
                  const int64_t *poff, int32_t ncurves, int64_t nptot,
                  const double *widths, int32_t nw, int32_t nl, double dtep, int32_t stat,
                  void *ws, int64_t ws_bytes, int32_t nblocks, void *stream);

/*
 * Execution geometry of the tls_fold launches that follow (no reference
 * counterpart; the autotuner's search space): epochs per staged LDS tile
 * (1..144, default 144), phase-index cells per staged sample (1 or 2,
 * default 2), samples per phase bucket of the counting sort (1..8, default
 * 2) and epochs per window-table run (0 = as many as the LDS holds, at most
 * 48).  Region membership and the best trial do not depend on it; the
 * cells and the table run change no bit, the tile epochs and the bucket size
 * (which set the staged sample ranges and the tiles' reference epochs) move
 * S/N by float64 rounding only.  Process
 * wide, read at each launch (a captured graph keeps the values of its
 * capture).
 */
int gtrap_tlsfold_set_tiles(int32_t max_tile_epochs, int32_t cells_per_sample,
                            int32_t samples_per_bucket, int32_t table_run_epochs);

/*
 * gtrap_tlsfold with no allocation and no workspace argument.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The work ticket and the per-workgroup
 * slots live in a static 64 MiB device array per GPU (held from the
 * library's load until the process ends); it calls no hipMalloc/hipFree.
 * nblocks is clamped to the slots that fit: 64 MiB / gtrap_tlsfold_ws_bytes
 * (nmax, 1), e.g. 30 workgroups for nmax = 70000, 511 for nmax = 4096; an
 * nmax whose single slot does not fit (about 3.3 million samples) is
 * rejected with GTRAP_EINVAL.  Results are identical to gtrap_tlsfold's.
 * Calls that overlap on one GPU (other streams or host threads) share the
 * array and race: serialise them.  The ticket is reset by a kernel on the
 * call's stream before every launch.
 */
int gtrap_tlsfold_static(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                         int32_t *best_e, const float *t, const float *x, const int64_t *off,
                         const int32_t *len, int32_t nmax, const double *periods,
                         const int64_t *poff, int32_t ncurves, int64_t nptot,
                         const double *widths, int32_t nw, int32_t nl, double dtep,
                         int32_t stat, int32_t nblocks, void *stream);

/*
 * Precision study of gtrap_tlsfold (no reference counterpart): the same
 * search with one arithmetic choice changed, for measuring single-precision
 * designs against the float64 kernel.  variant: 0 = gtrap_tlsfold itself;
 * bits 1 = flux stored as binary16 (x16, with x = h * x16scale[c] after the
 * curve's offset was removed), 2 = phase folded in float32 (quotient,
 * product and subtraction in float32), 4 = tile prefix sums accumulated in
 * float32, 8 = the same with two-sum compensation; the combinations 1, 2, 4,
 * 8, 1|8 and 4|2 are built.  Other arguments as gtrap_tlsfold (x may be NULL
 * in the binary16 variants).
 */
int gtrap_tlsfold_prec(int32_t variant, double *snr, double *hbest, int32_t *best_w,
                       int32_t *best_l, int32_t *best_e, const float *t, const float *x,
                       const uint16_t *x16, const float *x16scale, const int64_t *off,
                       const int32_t *len, int32_t nmax, const double *periods,
                       const int64_t *poff, int32_t ncurves, int64_t nptot,
                       const double *widths, int32_t nw, int32_t nl, double dtep, int32_t stat,
                       void *ws, int64_t ws_bytes, int32_t nblocks, void *stream);

/*
 * Item-list form of gtrap_tlsfold for the serving loop (no reference
 * counterpart).  Work item k scores period item_pidx[k] (index into periods
 * and the outputs) of curve slot item_curve[k] (off/len tables); the item
 * count is *nitems_dev and items are handed out through the atomic ticket
 * *counter, which the caller zeroes before the launch.  nblocks persistent
 * workgroups; launch shape fixed (HIP-graph capturable).  Same sample rules
 * and len[c] <= nmax precondition as gtrap_tlsfold.
 */
int gtrap_tlsfold_items(double *snr, double *hbest, int32_t *best_w, int32_t *best_l,
                        int32_t *best_e, const float *t, const float *x, const int64_t *off,
                        const int32_t *len, int32_t nmax, const double *periods,
                        const int32_t *item_curve, const int64_t *item_pidx,
                        const int32_t *nitems_dev, int32_t *counter, const double *widths,
                        int32_t nw, int32_t nl, double dtep, int32_t stat, void *ws,
                        int64_t ws_bytes, int32_t nblocks, void *stream);

/*
 * Periodogram statistics of the tls_fold output, the TLS counterpart of
 * gtrap/getstat.py:100-123: for curve slots curves[0 .. *ncurves_dev) (at most
 * max_curves), over periods [poff[c], poff[c] + np[c]): best_pidx[c] = index
 * of the first maximum S/N over the values that are neither NaN nor -inf
 * (-inf: a period without any trial; +inf, a perfect fit, can win; -1 if
 * none), best_sn[c], and sde[c] = (best - mean)/std with the mean and
 * standard deviation over the finite S/N values.  float64.
 */
int gtrap_tlsstats(const double *sn, const int64_t *poff, const int32_t *np_,
                   const int32_t *curves, const int32_t *ncurves_dev, int32_t max_curves,
                   int64_t *best_pidx, double *best_sn, double *sde, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GTRAP_HIP_H */
