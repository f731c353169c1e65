// Periodic trapezoid least squares (tls_fold): the kernel template and its
// device helpers, shared by tlsfold.hip (the production instantiations and
// the entry points) and tlsfold_prec.hip (the precision-study variants).
// The design notes are at the top of tlsfold.hip.
#pragma once
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kTfThreads = 512;
constexpr int kTfChunk = 2560;      // samples staged in LDS per tile/chunk (~123 KB)
constexpr int kTfMaxNl = 12;
constexpr int kTfMaxNw = 16;
constexpr int kTfMaxBnd = 3 + 4 * kTfMaxNl;  // boundaries per W

// Arithmetic variants of the precision study (template bits; 0 is the
// production kernel: float32 flux, float64 fold, float64 prefix sums).
constexpr int kVarF16 = 1;     // flux stored as binary16 after offset and power-of-2 scale
constexpr int kVarFold32 = 2;  // phase folded in float32 (quotient, product, subtraction)
constexpr int kVarAcc32 = 4;   // tile prefix sums accumulated in float32
constexpr int kVarAccTS = 8;   // tile prefix sums in float32 with two-sum compensation

// fmod(t, P) for t >= 0, P > 0, exactly: the remainder t - q*P of the
// integer quotient q is representable, so one fused multiply-add returns it
// exactly.  q = floor(t * invP) (invP = 1/P rounded) is within one of
// floor(t/P) (absolute error below 2^-2 for quotients below 2^50), and one
// correction step makes it exact.
__device__ __forceinline__ double tf_fmod(double t, double P, double invP) {
  double q = floor(t * invP);
  double r = fma(-q, P, t);
  if (r < 0.0) { q -= 1.0; r = fma(-q, P, t); }
  else if (r >= P) { q += 1.0; r = fma(-q, P, t); }
  return r;
}

// The fold of a variant: tf_fmod in float64 (production), or the float32
// arithmetic a single-precision implementation would use: P rounded to
// float32, the quotient floor(t/P) in float32, the product rounded before the
// subtraction, the result moved into [0, P).
template <int kVar>
__device__ __forceinline__ double tf_fold(float t, double P, double invP) {
  if constexpr ((kVar & kVarFold32) != 0) {
    const float Pf = (float)P;
    const float q = floorf(t / Pf);
    float r = t - __fmul_rn(q, Pf);
    if (r < 0.0f) r += Pf;
    else if (r >= Pf) r -= Pf;
    return (double)r;
  } else {
    return tf_fmod((double)t, P, invP);
  }
}

// A value every lane of the wave holds (read from lane 0): the compiler then
// keeps it in a scalar register and branches on it with scalar branches.
__device__ __forceinline__ int tf_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t tf_uniform64(int64_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((unsigned long long)hi << 32) | lo);
}

// Total order on float bit patterns (signed-magnitude to two's complement);
// breaks ties between samples of equal phase deterministically, NaN included.
__device__ __forceinline__ int tf_xkey(float x) {
  const int b = __float_as_int(x);
  return b ^ ((b >> 31) & 0x7fffffff);
}

// A sample takes part when its time is valid (t >= 0; masked samples carry a
// negative time) and its flux is finite; other samples are skipped.
__device__ __forceinline__ bool tf_usable(float t, float x) {
  return t >= 0.0f && isfinite(x);
}

// A trial's rank key is k = sign(S/N) * (S/N)^2, a monotone function of the
// S/N, so trials compare by key; the S/N and H of a thread's best trial come
// from its key and its numerator and denominator of H (tf_sn, tf_h).
// A trial's (w, l, e) is packed as ((w nl + l) << 32 | e) + 1, whose integer
// order is the (w, l, e) order; 0 means "none yet".
struct TfBest {
  double key, numB, denA;
  unsigned long long idx;
};
__device__ __forceinline__ unsigned long long tf_idx(int w, int l, int e, int nl) {
  return ((unsigned long long)(unsigned)(w * nl + l) << 32 | (unsigned)e) + 1ull;
}

// better: larger key, ties to the earliest (w, l, e).  Against "none yet"
// (key -inf, idx 0) every key but NaN and -inf is better, as in the oracle.
__device__ __forceinline__ bool tf_better(double key, unsigned long long idx, const TfBest &b) {
  return key > b.key || (key == b.key && idx < b.idx);
}

// Region boundary offsets of width W and step m (tls.py:46-50):
// kind 0: first index with d >= c ; kind 1: first index with d > c.
__device__ __forceinline__ void tf_bounds(double W, int m, int nl, double &tAB, double &tBC,
                                          double &tCD, double &tDE) {
  const double dl = 0.45 * W / (double)nl;
  const double L = (double)m * dl;
  const double qw = 0.25 * W, hl = 0.5 * L;
  tAB = -qw - hl;
  tBC = -qw + hl;
  tCD = qw - hl;
  tDE = qw + hl;
}

// One sample of the phase-sorted curve in the workspace: phase u, flux x.
// Sixteen bytes, so a sample moves in one store or load.
struct __align__(16) TfRec {
  double u;
  float x;
  int pad;
};

// (u, flux bits) order of two records
__device__ __forceinline__ bool tf_rec_less(const TfRec &a, const TfRec &b) {
  return a.u < b.u || (a.u == b.u && tf_xkey(a.x) < tf_xkey(b.x));
}

// Sorts rec[0, m) by (u, flux bits) with the whole workgroup: a bitonic network
// in its all-ascending form (every merge opens with a flip step), over the
// next power of two N >= m with virtual +inf elements at m..N-1, so a
// compare-exchange whose partner lies past m leaves its element in place.
// O(m log^2 m) compare-exchanges on the workspace; it serves phase buckets
// too large for one LDS chunk (strongly clustered phases, e.g. repeated
// time stamps) and curves with more LDS chunks than the chunk table holds.
__device__ void tf_bitonic(TfRec *rec, int m) {
  int N = 1;
  while (N < m) N <<= 1;
  for (int k = 2; k <= N; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = threadIdx.x; p < (N >> 1); p += kTfThreads) {
        // pair p -> its lower element i (a zero bit inserted at j's position)
        const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
        const int q = j == (k >> 1) ? (i ^ (k - 1)) : (i + j);
        if (q < m) {
          const TfRec ri = rec[i], rq = rec[q];
          if (tf_rec_less(rq, ri)) {
            rec[i] = rq;
            rec[q] = ri;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Extended index j of the sorted folded curve (nv samples) maps to sample
// tf_src(nv, j); its image is u - P for j < 0, u + P for j >= nv, else u.
__device__ __forceinline__ int tf_src(int nv, int j) {
  return j < 0 ? j + nv : (j >= nv ? j - nv : j);
}

// first index k in [lo, hi) of the LDS images with lu[k] - t0 >= c (kind 0)
// or > c (kind 1)
__device__ __forceinline__ int tf_search(const double *lu, int lo, int hi, double t0, double c,
                                         int kind) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    const double d = lu[mid] - t0;
    const bool pred = kind ? (d > c) : (d >= c);
    if (pred) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// Conservative extended index bound of value v from the bucket offsets
// (end[b] = cursor[b] after the scatter): upper=0 -> an index <= the first
// image >= v; upper=1 -> an index >= the last image <= v, plus one.
__device__ __forceinline__ int tf_bucket_bound(const int *cursor, int nb, int nv, double P,
                                               double bscale, double v, int upper) {
  int sh = 0;
  double w = v;
  if (v < 0.0) { sh = -1; w = v + P; }
  else if (v >= P) { sh = 1; w = v - P; }
  int bb = (int)(w * bscale);
  bb = bb < 0 ? 0 : (bb > nb - 1 ? nb - 1 : bb);
  int idx;
  if (!upper) {
    const int b0 = bb - 1;
    idx = b0 < 0 ? 0 : cursor[b0];  // start of bucket bb-1
  } else {
    const int b1 = bb + 1 > nb - 1 ? nb - 1 : bb + 1;
    idx = cursor[b1];  // end of bucket bb+1
  }
  idx += sh * nv;
  return idx < -nv ? -nv : (idx > 2 * nv ? 2 * nv : idx);
}

// Region sums of one trial: counts n*, flux sums s* (x' = x - x_ref), and for
// the ingress/egress regions B and D the sums of d, d^2 and x'd (d = offset
// from the epoch); the window's mean of x' over [-W/2, W/2] and its sum of
// (x' - mean)^2 over [-W/2, W/2) (both per (epoch, W), shared by every L).
struct TfRegions {
  double nA, nB, nC, nD, nE, sA, sB, sC, sD, sE;
  double dB, dD, d2B, d2D, xdB, xdD;
  double mean, y2;
};

// Window statistics of an (epoch, W): mean over the closed window and the
// sum of squares of x' - mean over [-W/2, W/2) (tls.py:57-59 and the residual's
// constant term).
__device__ __forceinline__ void tf_window(double mcnt, double sM, double nAll, double s1All,
                                          double s2All, double &mean, double &y2) {
  mean = sM / mcnt;
  y2 = s2All - 2.0 * mean * s1All + mean * mean * nAll;
}

// The statistic of one trial from its region sums (tls.py:44-79 expanded):
// H = -numB/denA and the residual res = y2 + H U + H^2 V expanded in the
// sums, so res denA^2 = y2 denA^2 - numB denA U + numB^2 V, and
// S/N = -H / sqrt(res/(n-3)) has the rank key
// k = sign(numB denA) numB^2 (n-3) / (res denA^2): one division per trial.
// With stat = 1 the trial is the least-squares fit of the same trapezoid
// shape instead (the detection statistic): res(H) = y2 + H U + H^2 V is
// minimal at H* = -U/(2V), where the chi^2 improvement is U^2/(4V); the key
// is sign(-H*) U^2/(4V) = sign(U) U^2/(4V) (no division by the window's own
// residual: the reported value divides sqrt(|key|) by the curve's standard
// deviation), and numB and denA are U and 2V (H = -numB/denA = H*).
// c is the trial's (W, L) table row (products of L and W at 4..9).  Writes
// numB and denA; NaN for no trial (n <= 3, denA = 0).  Multiply-adds may
// fuse here: the statistic's rounding is not part of any region membership.
__device__ __forceinline__ double tf_eval(const TfRegions &g, double W, const double *c,
                                          int stat, double *numBout, double *denAout) {
#pragma clang fp contract(fast)
  const double c4L2 = c[4], c8L2 = c[5], c4WL = c[6], c16L = c[7], c2iL = c[8], ciL2 = c[9];
  const double q4 = 0.25 * W, q42 = q4 * q4, W2 = W * W, W8 = 8.0 * W;
  const double mean = g.mean;
  const double yA = g.sA - mean * g.nA, yB = g.sB - mean * g.nB, yC = g.sC - mean * g.nC,
               yD = g.sD - mean * g.nD, yE = g.sE - mean * g.nE;
  const double ydB = g.xdB - mean * g.dB, ydD = g.xdD - mean * g.dD;
  const double nAE = g.nA + g.nE, nBD = g.nB + g.nD, yAE = yA + yE;
  const double denA =
      c4L2 * (nAE - g.nC) + W2 * nBD + W8 * (g.dB - g.dD) + 16.0 * (g.d2B + g.d2D);
  const double numB = c8L2 * (yAE + yC) - c4WL * (yB + yD) + c16L * (ydD - ydB);
  // res = y2 + 2h(yA - yC + yE) + h^2(nA + nC + nE) - 2a(ydB + q4 yB)
  //       + a^2(d2B + 2 q4 dB + q4^2 nB) + 2a(ydD - q4 yD) + a^2(d2D - 2 q4 dD + q4^2 nD),
  // h = H/2, a = H/L, q4 = W/4
  const double U = (yAE - yC) + c2iL * ((ydD - q4 * yD) - (ydB + q4 * yB));
  const double V = 0.25 * (nAE + g.nC) +
                   ciL2 * ((g.d2B + 2.0 * q4 * g.dB + q42 * g.nB) +
                           (g.d2D - 2.0 * q4 * g.dD + q42 * g.nD));
  const double R = g.y2 * denA * denA - numB * denA * U + numB * numB * V;
  const double nn = nAE + nBD + g.nC;  // integers: exact in any order
  *numBout = numB;
  *denAout = denA;
  constexpr double kNaN = __builtin_nan("");
  if (stat) {
    // least-squares matched filter of the trapezoid shape: signed chi^2
    // improvement (the division by 4V is the same for every trial's rank
    // only up to V, so it stays)
    const double k = U * U / (4.0 * V);
    double key = U < 0.0 ? -k : k;
    if (nn <= 3.0 || V == 0.0) key = kNaN;
    *numBout = U;
    *denAout = 2.0 * V;
    return key;
  }
  // (S/N)^2 = H^2 (n-3)/res = numB^2 (n-3)/R
  const double s2 = numB * numB * (nn - 3.0);
  const bool neg = (numB < 0.0) != (denA < 0.0);
  const double k = s2 / R;  // every case below is a select (no branch)
  double key = neg ? -k : k;
  // the expanded residual of a perfect fit can round below zero: a residual
  // of 0 gives an infinite S/N (0/0, NaN, when H = 0 too)
  if (R <= 0.0) key = numB == 0.0 ? kNaN : (neg ? -INFINITY : INFINITY);
  // n - 3 degrees of freedom: fewer than 4 samples is no trial (the direct
  // formula gives 0/0 there); denA = 0 makes H infinite and the S/N NaN
  if (nn <= 3.0 || denA == 0.0) key = kNaN;
  return key;
}

// S/N and H of a best trial from its key and H's numerator and denominator
__device__ __forceinline__ double tf_sn(double key) {
  return key >= 0.0 ? sqrt(key) : -sqrt(-key);
}
__device__ __forceinline__ double tf_h(double numB, double denA) { return -numB / denA; }

struct TfArgs {
  const float *t, *x;
  const int64_t *off;
  const int32_t *len;
  const double *periods;
  const int64_t *poff;
  const int32_t *np;
  int32_t ncurves;
  int64_t nptot;
  const double *widths;
  int32_t nw, nl;
  double dtep;
  double *sn, *hbest;
  int32_t *bw, *bl, *be;
  unsigned char *ws;
  int64_t ws_per_block;
  int32_t nmax;
  // item-list mode (serving): item k is period item_pidx[k] of curve slot
  // item_curve[k]; the count is read from *nitems_dev.  counter != null
  // hands items out dynamically (atomic ticket), else grid-stride.
  const int32_t *item_curve;
  const int64_t *item_pidx;
  const int32_t *nitems_dev;
  int32_t *counter;
  int32_t te_run;  // epochs per window-table run (fits the dynamic LDS)
  int32_t stat;    // 0: tls.py's S/N; 1: the chi^2-improvement detection statistic
  // execution geometry (results do not depend on it beyond float64 rounding)
  int32_t max_tile;   // epochs per staged tile (at most kTfMaxTile)
  int32_t cells;      // phase-index cells per staged sample (1 or 2)
  int32_t spb;        // samples per phase bucket of the fold's counting sort (1..8)
  // binary16 flux of the kVarF16 variants: x = h * x16scale[c] (offset removed)
  const unsigned short *x16;
  const float *x16scale;
};

// LDS layout for a staged range
constexpr int kTfMaxTe = 48;     // epochs per window-table run (at most)
constexpr int kTfWl = 10;  // (W, L) table row: tAB, tBC, tCD, tDE, 4L^2, 8L^2, 4WL, 16L,
                           // 2/L, 1/L^2
constexpr int kTfCells = 2 * kTfChunk;  // phase index cells of a staged range (at most)
constexpr int kTfMaxTile = 144;  // epochs per staged tile: with the reference at the
                                 // tile centre, |q| stays within 72 epoch steps + W/2
constexpr int kTfMaxBuckets = 30000;  // phase buckets (LDS counters, <= 120 KB)
constexpr int kTfSortChunk = 7680;    // samples per in-LDS bucket-sort chunk (90 KB + ends)
constexpr int kTfMaxBreaks = 128;     // sort chunks per period (longer curves: one bitonic sort)
constexpr int kTfUnroll = 16;         // global loads in flight per thread in the sort loops (8: 0.25% slower)

// Static LDS of a workgroup.  The per-(W, boundary) accumulators of the chunked
// sweep and the per-(epoch, W) window table are sized at launch (dynamic LDS).
struct TfLds {
  // staged images (exact), then 8 entries of +inf (read by the bisections)
  double lu[kTfChunk + 8];
  double px[kTfChunk + 1][5];   // exclusive prefix of x', x'^2, q, q^2, x'q
  // phase index of the staged range (two cells per sample): T(g) = first k
  // with cell(lu[k]) >= g, cell(v) = clamp(floor((v - tbase) * tinv), 0, tnc - 1),
  // for g = -2 .. tnc + 2 (T(-2) = T(-1) = 0, T(tnc) .. T(tnc + 2) = cnt),
  // stored as tab[g + 2]
  unsigned short tab[kTfCells + 5];
  double tbase, tinv;
  int tnc;
  double scan[kTfThreads / kWave][5];
  double red_key[kTfThreads / kWave], red_nb[kTfThreads / kWave], red_da[kTfThreads / kWave];
  unsigned long long red_idx[kTfThreads / kWave];
};

// boundary b of width W (kind, offset) from the width's row of the (W, L)
// table: 0 -> -W/2 (>=), 1 -> W/2 (>=, region end), 2 -> W/2 (>, mean end),
// 3 + 4(m-1) + {0,1,2,3} -> tAB, tBC, tCD, tDE of step m >= 1 (>=)
__device__ __forceinline__ void tf_bnd(double W, const double *wrow, int b, double &c,
                                       int &kind) {
  if (b == 0) { c = -0.5 * W; kind = 0; return; }
  if (b == 1) { c = 0.5 * W; kind = 0; return; }
  if (b == 2) { c = 0.5 * W; kind = 1; return; }
  const int m = ((b - 3) >> 2) + 1, r = (b - 3) & 3;
  c = wrow[kTfWl * m + r];
  kind = 0;
}

// cell of v in the phase index of a staged range
__device__ __forceinline__ int tf_cell(double v, double base, double inv, int nc) {
  return (int)fmin(fmax((v - base) * inv, 0.0), (double)(nc - 1));  // NaN -> 0
}

// Lane move of a double by a DPP control (both 32-bit halves).  Lanes whose
// source lies outside their row, and lanes of rows outside ROWS, read +0.0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double tf_dpp0(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Inclusive scans over the 64 lanes of a wave of N values at once, with DPP
// moves (no LDS crossbar): row_shr 1, 2, 4, 8 inside each row of 16 lanes,
// then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.  A
// lane without a source adds +0.0, which leaves every sum as it is (but for
// the sign of a zero).
template <int CTRL, int ROWS, int N>
__device__ __forceinline__ void tf_scan_step(double (&v)[N]) {
  double t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = tf_dpp0<CTRL, ROWS>(v[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += t[i];
}
template <int N>
__device__ __forceinline__ void tf_wave_iscan(double (&v)[N]) {
  tf_scan_step<0x111, 0xf>(v);
  tf_scan_step<0x112, 0xf>(v);
  tf_scan_step<0x114, 0xf>(v);
  tf_scan_step<0x118, 0xf>(v);
  tf_scan_step<0x142, 0xa>(v);
  tf_scan_step<0x143, 0xc>(v);
}

// Stage extended range [lo, lo+cnt) (cnt <= kTfChunk) into LDS: the images
// lu, the phase index and the exclusive prefix sums of (x', x'^2, q, q^2,
// x'q), q = image - r.  Thread t owns the kPer consecutive elements from
// k0 = t kPer and reads them (and the image before them, for the phase index)
// straight from the records, so one barrier (after the wave totals) and the
// closing one are all it needs.
constexpr int kPer = kTfChunk / kTfThreads;

// float32 pair accumulated with two-sum (the kVarAccTS variants): hi holds
// the rounded sum, lo the accumulated rounding errors
struct TfF2 {
  float hi, lo;
};
__device__ __forceinline__ TfF2 tf_f2_add(TfF2 a, float b) {
  const float s = a.hi + b;
  const float bb = s - a.hi;
  const float e = (a.hi - (s - bb)) + (b - bb);
  return TfF2{s, a.lo + e};
}
__device__ __forceinline__ TfF2 tf_f2_add2(TfF2 a, TfF2 b) {
  TfF2 r = tf_f2_add(a, b.hi);
  r.lo += b.lo;
  return r;
}

// The prefix sums of a staged range in float32 (kVarAcc32) or float32
// two-sum (kVarAccTS) arithmetic, from the float64 per-sample values xs, qs
// rounded to float32; written to L.px as float64.  Same layout and
// synchronisation as the float64 path of tf_stage.
template <int kVar>
__device__ void tf_stage_prefix32(TfLds &L, const double *xs, const double *qs, int k0, int cnt) {
  constexpr bool kTS = (kVar & kVarAccTS) != 0;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  float xf[kPer], qf[kPer];
  TfF2 acc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) acc[v] = TfF2{0.0f, 0.0f};
  auto add = [&](TfF2 a, float b) { return kTS ? tf_f2_add(a, b) : TfF2{a.hi + b, 0.0f}; };
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    xf[u] = (float)xs[u];
    qf[u] = (float)qs[u];
    if (k0 + u < cnt) {
      acc[0] = add(acc[0], xf[u]);
      acc[1] = add(acc[1], __fmul_rn(xf[u], xf[u]));
      acc[2] = add(acc[2], qf[u]);
      acc[3] = add(acc[3], __fmul_rn(qf[u], qf[u]));
      acc[4] = add(acc[4], __fmul_rn(xf[u], qf[u]));
    }
  }
  TfF2 inc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) inc[v] = acc[v];
  for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const TfF2 t{__shfl_up(inc[v].hi, o), __shfl_up(inc[v].lo, o)};
      if (lane >= o) inc[v] = kTS ? tf_f2_add2(t, inc[v]) : TfF2{t.hi + inc[v].hi, 0.0f};
    }
  }
  if (lane == kWave - 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) L.scan[wv][v] = (double)inc[v].hi + (double)inc[v].lo;
  }
  __syncthreads();
  TfF2 run[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    TfF2 base{0.0f, 0.0f};
    for (int w = 0; w < wv; ++w) {
      const double d = L.scan[w][v];
      const float hi = (float)d;
      base = kTS ? tf_f2_add2(base, TfF2{hi, (float)(d - (double)hi)}) : TfF2{base.hi + hi, 0.0f};
    }
    // exclusive prefix: the block base plus the lanes before this one
    const TfF2 before{__shfl_up(inc[v].hi, 1), __shfl_up(inc[v].lo, 1)};
    run[v] = lane == 0 ? base : (kTS ? tf_f2_add2(base, before) : TfF2{base.hi + before.hi, 0.0f});
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = k0 + u;
    if (k < cnt) {
#pragma unroll
      for (int v = 0; v < 5; ++v) L.px[k][v] = (double)run[v].hi + (double)run[v].lo;
      run[0] = add(run[0], xf[u]);
      run[1] = add(run[1], __fmul_rn(xf[u], xf[u]));
      run[2] = add(run[2], qf[u]);
      run[3] = add(run[3], __fmul_rn(qf[u], qf[u]));
      run[4] = add(run[4], __fmul_rn(xf[u], qf[u]));
      if (k == cnt - 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.px[cnt][v] = (double)run[v].hi + (double)run[v].lo;
      }
    }
  }
  if (cnt <= 0 && tid == 0) {
#pragma unroll
    for (int v = 0; v < 5; ++v) L.px[0][v] = 0.0;
  }
  __syncthreads();
}

template <int kVar>
__device__ void tf_stage(TfLds &L, const TfRec *rec, int nv, double P, int lo,
                         int cnt, double r, double xref, int cells) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int k0 = tid * kPer;
  // image of extended element j = lo + k of record value u: u - P, u or
  // u + P, as one fma with the shift -1, 0 or +1 (rounded exactly as the sum;
  // a phase of -0 would come out +0, which no comparison or sum tells apart);
  // selecting between both sums cost 1.2% more on the serving mix
  auto image = [&](int k, double u) {
    const int j = lo + k;
    return fma((double)((j >= nv) - (j < 0)), P, u);
  };
  double img[kPer], xs[kPer], qs[kPer];
  double acc[5] = {0, 0, 0, 0, 0};
  if (cnt > 0) {
    // every load in flight before the first use: the range ends (the same two
    // records for every thread), the element before k0 and the own elements
    const int last = cnt - 1;
    const double ub = rec[tf_src(nv, lo)].u, ut = rec[tf_src(nv, lo + last)].u;
    const double up = rec[tf_src(nv, lo + min(max(k0 - 1, 0), last))].u;
    TfRec rv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) rv[u] = rec[tf_src(nv, lo + min(k0 + u, last))];
    const double base = image(0, ub), top = image(last, ut);
    // phase index: two cells per sample; cells narrower than 1e-9 of the
    // magnitudes (a range of near-equal phases) fall back to one cell
    const int nc = cells * min(cnt, kTfChunk);
    const double span = top - base;
    const double inv =
        span > 1e-9 * (1.0 + fabs(base) + fabs(top)) * (double)nc ? (double)nc / span : 0.0;
    int cp = k0 == 0 ? -1 : tf_cell(image(k0 - 1, up), base, inv, nc);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = k0 + u;
      img[u] = image(k, rv[u].u);
      xs[u] = 0.0;
      qs[u] = 0.0;
      if (k < cnt) {
        L.lu[k] = img[u];
        xs[u] = (double)rv[u].x - xref;
        qs[u] = img[u] - r;
        acc[0] += xs[u]; acc[1] += xs[u] * xs[u]; acc[2] += qs[u]; acc[3] += qs[u] * qs[u];
        acc[4] += xs[u] * qs[u];
        // cells (cp, ck] start at k: the first two by predicated stores,
        // longer runs (gaps in phase) by a loop
        const int ck = tf_cell(img[u], base, inv, nc);
        if (ck > cp) L.tab[cp + 3] = (unsigned short)k;
        if (ck > cp + 1) L.tab[cp + 4] = (unsigned short)k;
        for (int g = cp + 3; g <= ck; ++g) L.tab[g + 2] = (unsigned short)k;
        cp = ck;
      }
    }
    const int cl = tf_cell(top, base, inv, nc);
    for (int g = cl + 1 + tid; g <= nc + 2; g += kTfThreads) L.tab[g + 2] = (unsigned short)cnt;
    if (tid == 0) { L.tab[0] = 0; L.tab[1] = 0; L.tbase = base; L.tinv = inv; L.tnc = nc; }
  } else if (tid == 0) {
    for (int g = 0; g < 6; ++g) L.tab[g] = 0;
    L.tbase = 0.0;
    L.tinv = 0.0;
    L.tnc = 1;
  }
  if (tid < 8) L.lu[max(cnt, 0) + tid] = INFINITY;
  if constexpr ((kVar & (kVarAcc32 | kVarAccTS)) != 0) {
    tf_stage_prefix32<kVar>(L, xs, qs, k0, cnt);
    return;
  }
  // inclusive wave scan of the per-thread totals
  double inc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) inc[v] = acc[v];
  tf_wave_iscan(inc);
  if (lane == kWave - 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) L.scan[wv][v] = inc[v];
  }
  __syncthreads();
  // the totals of the waves before this one, added in wave order; every
  // wave's total is read at once (a loop to wv would chain up to 35 LDS round
  // trips), the select keeps the sum's order and bits
  double run[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double tot[kTfThreads / kWave];
#pragma unroll
    for (int w = 0; w < kTfThreads / kWave; ++w) tot[w] = L.scan[w][v];
    double base = 0.0;
#pragma unroll
    for (int w = 0; w < kTfThreads / kWave; ++w) base = w < wv ? base + tot[w] : base;
    run[v] = base + inc[v] - acc[v];  // exclusive prefix of this thread
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int k = k0 + u;
    if (k < cnt) {
#pragma unroll
      for (int v = 0; v < 5; ++v) L.px[k][v] = run[v];
      run[0] += xs[u]; run[1] += xs[u] * xs[u]; run[2] += qs[u]; run[3] += qs[u] * qs[u];
      run[4] += xs[u] * qs[u];
      if (k == cnt - 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.px[cnt][v] = run[v];
      }
    }
  }
  if (cnt <= 0 && tid == 0) {
#pragma unroll
    for (int v = 0; v < 5; ++v) L.px[0][v] = 0.0;
  }
  __syncthreads();
}

// prefix row p (p <= kTfChunk, so a 24-bit multiply forms its offset)
__device__ __forceinline__ const double *tf_px(const TfLds &L, int p) {
  return &L.px[0][0] + __umul24((unsigned)p, 5u);
}

// tf_search over the staged range through the phase index, f = (t0 + c -
// tbase) * tinv the target's fractional cell.  Cells are at least 1e-9 of the
// magnitudes wide (tf_stage), so the rounding of f and of the offset test is
// below 1e-4 of a cell: when f lies at least 1e-4 of a cell inside cell g,
// every sample of a lower cell fails the test and every sample of a higher
// cell passes it, and the answer lies in [T(g), T(g+1)].  Otherwise the
// answer lies in [T(g-1), T(g+2)] (one whole cell of margin each side).
// The result is clamped to [lo, hi], which equals the search restricted to
// [lo, hi) (the test is monotone in the sorted images).
// Both cases are index selects (no branch).  gi is the floor of f clamped to
// [-1, tnc] (tf_cell(t0 + c) for gi inside [0, tnc)).  Outside, both
// brackets still hold the answer: f below 0 (gi = -1) lies below every
// sample's cell coordinate, so every sample passes and the answer is 0 =
// T(-1) = T(0); f at or above tnc (gi = tnc) lies at or above every sample's
// (the top sample's is tnc), so with the margin every sample fails and the
// answer is cnt = T(tnc) = T(tnc + 1); near 0 or tnc the wide bracket
// [T(gi-1), T(gi+2)] covers both outcomes.  T's padding covers gi - 1 >= -2
// and gi + 2 <= tnc + 2.
constexpr double kTfEdge = 1e-4;
__device__ __forceinline__ void tf_bracket_f(const TfLds &L, double f, int &a, int &b) {
  const int nc = L.tnc;
  const double fl = floor(f), fr = f - fl;
  const int gi = (int)fmin(fmax(fl, -1.0), (double)nc);  // NaN -> -1
  const bool inner = fabs(fr - 0.5) <= 0.5 - kTfEdge;
  const int in = inner ? 1 : 0;
  a = L.tab[gi + 1 + in];  // T(gi - 1 + in)
  b = L.tab[gi + 4 - in];  // T(gi + 2 - in)
}
__device__ __forceinline__ void tf_bracket(const TfLds &L, double t0, double c, int &a, int &b) {
  tf_bracket_f(L, (t0 + c - L.tbase) * L.tinv, a, b);
}

// r clamped to [lo, hi] (lo <= hi)
__device__ __forceinline__ int tf_clamp(int r, int lo, int hi) { return min(max(r, lo), hi); }

// First index passing the offset test (kind 0, d >= c) in each of four
// brackets [a[i], b[i]) of at most kTfCount samples, by a branch-free
// bisection of fixed depth (steps 4, 2, 1) with the four searches in step:
// [a[i], r[i]) holds only failing samples throughout, and a step advances
// r[i] past failing samples only.  Samples at or past b[i] lie in later
// cells than the target (or are the +inf padding) and pass, so reads past the
// bracket need no mask.  Reads reach at most index a[i] + 6 <= cnt + 6,
// inside lu's padding.  The step of 4 tests sample a[i] + 3, which passes
// (leaving r[i] unchanged) when bracket i holds at most 3 samples; when all
// four do (wide = 0) the step is skipped, and a wave whose lanes all skip it
// saves a quarter of the search.
constexpr int kTfCount = 7;
__device__ __forceinline__ void tf_count4(const TfLds &L, const int *a, double t0,
                                          const double *c, bool wide, int *r) {
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = a[i];
  if (wide) {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] += (int)!(L.lu[r[i] + 3] - t0 >= c[i]) * 4;
  }
#pragma unroll
  for (int st = 2; st > 0; st >>= 1) {
    double d[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = r[i] + st - 1;
      d[i] = L.lu[k] - t0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int fail = (int)!(d[i] >= c[i]);
      r[i] += fail * st;
    }
  }
}

// One boundary of either kind (0: d >= c, 1: d > c) in a bracket [a, b) of
// at most kTfCount samples, by the same fixed-depth bisection as tf_count4
// (its step of 4 only for a bracket of more than 3 samples).
__device__ __forceinline__ int tf_bisect(const TfLds &L, int a, int b, double t0, double c,
                                         int kind) {
  int r = a;
  if (b - a > 3) {
    const double d = L.lu[a + 3] - t0;
    r += (int)!(kind ? d > c : d >= c) * 4;
  }
#pragma unroll
  for (int st = 2; st > 0; st >>= 1) {
    const double d = L.lu[r + st - 1] - t0;
    const int fail = (int)!(kind ? d > c : d >= c);
    r += fail * st;
  }
  return r;
}

// The statistic from the staged prefix sums at the trial's boundary indices:
// raw differences first (x', x'^2 for the window and regions; q, q^2, x'q
// only for B and D), shifted to epoch coordinates (d = q - s) per region.
// r0..rE are the prefix rows (x', x'^2, q, q^2, x'q) at the boundary indices.
__device__ __forceinline__ double tf_stat_rows(int p0, int pAB, int pBC, int pCD, int pDE,
                                               int pE, const double *r0, const double *rAB,
                                               const double *rBC, const double *rCD,
                                               const double *rDE, const double *rE, double mean,
                                               double y2, double s, double W, const double *c,
                                               int stat, double *numB, double *denA) {
#pragma clang fp contract(fast)
  const double x0 = r0[0], xAB = rAB[0], xBC = rBC[0], xCD = rCD[0], xDE = rDE[0], xE = rE[0];
  TfRegions g;
  g.mean = mean;
  g.y2 = y2;
  g.nA = (double)(pAB - p0); g.nB = (double)(pBC - pAB); g.nC = (double)(pCD - pBC);
  g.nD = (double)(pDE - pCD); g.nE = (double)(pE - pDE);
  g.sA = xAB - x0; g.sB = xBC - xAB; g.sC = xCD - xBC; g.sD = xDE - xCD; g.sE = xE - xDE;
  const double q1B = rBC[2] - rAB[2], q2B = rBC[3] - rAB[3], xqB = rBC[4] - rAB[4];
  const double q1D = rDE[2] - rCD[2], q2D = rDE[3] - rCD[3], xqD = rDE[4] - rCD[4];
  g.dB = q1B - g.nB * s;
  g.d2B = q2B - 2.0 * s * q1B + g.nB * s * s;
  g.xdB = xqB - s * g.sB;
  g.dD = q1D - g.nD * s;
  g.d2D = q2D - 2.0 * s * q1D + g.nD * s * s;
  g.xdD = xqD - s * g.sD;
  return tf_eval(g, W, c, stat, numB, denA);
}
__device__ __forceinline__ double tf_stat_px(const TfLds &L, int p0, int pAB, int pBC, int pCD,
                                             int pDE, int pE, double mean, double y2, double s,
                                             double W, const double *c, int stat, double *numB,
                                             double *denA) {
  return tf_stat_rows(p0, pAB, pBC, pCD, pDE, pE, tf_px(L, p0), tf_px(L, pAB), tf_px(L, pBC),
                      tf_px(L, pCD), tf_px(L, pDE), tf_px(L, pE), mean, y2, s, W, c, stat, numB,
                      denA);
}

// One boundary of a wide pass, resolved in the workspace: the prefix row
// (x', x'^2, q, q^2, x'q; cumulative from the pass's first sample, q relative
// to the pass's reference) at the boundary's index, and that index counted
// from the pass's first sample.
struct __align__(16) TfSlot {
  double row[5];
  int idx, pad;
};
// (epoch, W, boundary) slots of one wide pass: nmax/4 within [1024, 16384]
// (at least the 16 x 47 boundaries of one epoch at the largest nw and nl; a
// wide pass needs a curve of several thousand samples)
constexpr int kTfPassSlotsMin = 1024, kTfPassSlotsMax = 16384;
__host__ __device__ inline int tf_pass_slots(int64_t nmax) {
  const int64_t s = nmax / 4;
  return (int)(s < kTfPassSlotsMin ? kTfPassSlotsMin : (s > kTfPassSlotsMax ? kTfPassSlotsMax : s));
}

// Workspace slot of one workgroup: the phase-sorted records (nmax of 16
// bytes), the bucket ends (nmax + 2 ints), then, 16-byte aligned, the
// tf_pass_slots(nmax) slots of a wide pass.
__host__ __device__ inline int64_t tf_pass_off(int64_t nmax) {
  return (nmax * 16 + (nmax + 2) * 4 + 15) / 16 * 16;
}
__host__ __device__ inline int64_t tf_block_bytes(int64_t nmax) {
  return (tf_pass_off(nmax) + (int64_t)tf_pass_slots(nmax) * (int64_t)sizeof(TfSlot) + 255) /
         256 * 256;
}

static_assert(sizeof(TfLds) >= (kTfMaxBuckets + 1) * sizeof(int), "bucket counters exceed LDS");
static_assert(kTfPassSlotsMin >= kTfMaxNw * kTfMaxBnd, "a wide pass must hold one epoch");
static_assert(sizeof(TfLds) >= kTfSortChunk * (sizeof(double) + sizeof(float) + sizeof(int)),
              "sort chunk exceeds LDS");

// kItems: work items come from an item list (serving steps), else every
// period of a packed batch, longest first.  Two instantiations, so profiles
// tell the serving kernel from the per-request one.
template <bool kItems, int kVar>
__global__ __launch_bounds__(kTfThreads) void tlsfold_kernel(TfArgs A) {
  __shared__ TfLds L;
  __shared__ int s_cnt, s_i0, s_wsum[kTfThreads / kWave], s_brk[kTfMaxBreaks], s_nbrk;
  __shared__ int64_t s_item;
  __shared__ double s_sigma;  // the curve's flux standard deviation (stat 1)
  extern __shared__ __align__(16) unsigned char tf_dyn[];
  const int tid = threadIdx.x;
  const int nw = A.nw, nl = A.nl;
  // L = 0 (m = 0) empties the ingress/egress regions: denA = numB = 0 and the
  // trial's S/N is NaN, which never wins, so only m = 1..nl-1 are evaluated
  const int nlt = nl - 1;
  const int nbnd = 3 + 4 * nlt;
  // dynamic LDS: per-(W, L) region offsets and products of L, the widths, then the
  // per-(epoch, W) window table
  double *wlt = reinterpret_cast<double *>(tf_dyn);  // [nw*nl][kTfWl]
  double *sw = wlt + (size_t)kTfWl * nw * nl;        // [nw]
  double *ewd = sw + nw;                             // [te_run*nw][2]: window mean, y2
  int *ewp = reinterpret_cast<int *>(ewd + (size_t)2 * A.te_run * nw);  // [te_run*nw][3]
  for (int k = tid; k < nw * nl; k += kTfThreads) {
    const int w = k / nl, m = k % nl;
    const double W = A.widths[w];
    double *row = wlt + kTfWl * k;
    tf_bounds(W, m, nl, row[0], row[1], row[2], row[3]);
    const double Lm = (double)m * (0.45 * W / (double)nl);  // L = 0 rows are never used
    row[4] = 4.0 * Lm * Lm;
    row[5] = 8.0 * Lm * Lm;
    row[6] = 4.0 * W * Lm;
    row[7] = 16.0 * Lm;
    row[8] = 2.0 / Lm;
    row[9] = 1.0 / (Lm * Lm);
  }
  for (int w = tid; w < nw; w += kTfThreads) sw[w] = A.widths[w];
  __syncthreads();
  // this thread's first (epoch, W) entry of a window-table run
  const int wt_e = tid / nw, wt_w = tid % nw;
  // advance of the trial loop's (L, W, epoch) indices per kTfThreads trials
  const int dm = nlt ? kTfThreads % nlt : 0, dw = nlt ? (kTfThreads / nlt) % nw : 0,
            dje = nlt ? kTfThreads / (nlt * nw) : 0;
  unsigned char *ws = A.ws + 256 + (size_t)blockIdx.x * (size_t)A.ws_per_block;
  TfRec *rec = reinterpret_cast<TfRec *>(ws);
  int *cursor = reinterpret_cast<int *>(rec + A.nmax);
  TfSlot *slots = reinterpret_cast<TfSlot *>(ws + tf_pass_off(A.nmax));
  int *ticket = A.counter ? A.counter : reinterpret_cast<int *>(A.ws);
  double wmax = 0.0;
  for (int w = 0; w < nw; ++w) wmax = fmax(wmax, A.widths[w]);
  const double hwm = 0.5 * wmax;
  const double marg = 1e-9;
  const int64_t nitems = A.nitems_dev ? (int64_t)*A.nitems_dev : A.nptot;

  for (;;) {
    // ---- next work item (atomic ticket; packed batches run longest period first)
    if (tid == 0) s_item = (int64_t)atomicAdd(ticket, 1);
    __syncthreads();
    // The item's scalars are the same in every lane; read them as such
    // (readfirstlane), so that every branch on them is a scalar branch (the
    // compiler cannot tell on its own that values loaded through the LDS
    // ticket are uniform).
    const int64_t slot = tf_uniform64(s_item);
    __syncthreads();
    if (slot >= nitems) break;
    int c;
    int64_t item;
    if (kItems) {
      c = tf_uniform(A.item_curve[slot]);
      item = tf_uniform64(A.item_pidx[slot]);
    } else {
      item = A.nptot - 1 - slot;
      int lo_c = 0, hi_c = A.ncurves - 1;
      while (lo_c < hi_c) {
        const int mid = (lo_c + hi_c + 1) >> 1;
        if (A.poff[mid] <= item) lo_c = mid; else hi_c = mid - 1;
      }
      c = lo_c;
    }
    const double P = A.periods[item];
    const float *tc = A.t + A.off[c];
    const float *xc = A.x + A.off[c];
    // flux of sample i (binary16 times the curve's scale in the F16 variants)
    const unsigned short *hc = (kVar & kVarF16) ? A.x16 + A.off[c] : nullptr;
    const float hscale = (kVar & kVarF16) ? A.x16scale[c] : 1.0f;
    auto xload = [&](int i) -> float {
      if constexpr ((kVar & kVarF16) != 0) {
        return (float)__builtin_bit_cast(_Float16, hc[i]) * hscale;
      } else {
        return xc[i];
      }
    };
    const int n = tf_uniform(A.len[c]);
    const int nE = (int)ceil(P / A.dtep);
    if (n > A.nmax) {
      // a curve longer than the workspace slot is not searched (the host
      // entry points guard this where the lengths are known): S/N NaN and
      // indices -2 mark the item
      if (tid == 0) {
        A.sn[item] = NAN;
        if (A.hbest) A.hbest[item] = NAN;
        A.bw[item] = -2;
        A.bl[item] = -2;
        A.be[item] = -2;
      }
      // every lane meets here before the next ticket, as at the end of a
      // searched item: without this barrier the compiler joined thread 0's
      // writes to the next ticket's atomic (both under tid == 0), so thread 0
      // and the other lanes of its wave reached the loop's first barrier by
      // different paths and the wave never reconverged
      __syncthreads();
      continue;
    }

    // ---- 1. fold and bucket-sort by phase ------------------------------
    // Counting sort into ~2-sample phase buckets with LDS counters (the tile
    // LDS is idle during this phase), scatter to the workspace, then sort
    // every bucket by (u, x) in LDS in bucket-aligned chunks.  The bucket ends
    // are kept in the workspace for the tile phase.
    // n/2 buckets (counted over all samples, so the counting pass also finds
    // the number of valid samples and the first valid index)
    // (each block's workspace slot holds nmax/2 + 2 bucket ends after the
    // records, gtrap_tlsfold_ws_bytes: more buckets need a larger slot)
    const int nb = max(1, min(n / A.spb, kTfMaxBuckets));
    const double bscale = (double)nb / P;
    const double invP = 1.0 / P;
    int *lcur = reinterpret_cast<int *>(&L);  // nb + 1 <= kTfMaxBuckets + 1 ints
    if (tid == 0) { s_cnt = 0; s_i0 = n; }
    for (int b = tid; b <= nb; b += kTfThreads) lcur[b] = 0;
    __syncthreads();
    double mysx = 0.0, mysxx = 0.0;  // flux sums for the curve's deviation (stat 1)
    {
      int myv = 0, myi0 = n;
      for (int i0 = tid; i0 < n; i0 += kTfUnroll * kTfThreads) {
        float tv[kTfUnroll], xv[kTfUnroll];  // every load in flight before the first use
#pragma unroll
        for (int v = 0; v < kTfUnroll; ++v) {
          const int i = min(i0 + v * kTfThreads, n - 1);
          tv[v] = tc[i];
          xv[v] = xload(i);
        }
#pragma unroll
        for (int v = 0; v < kTfUnroll; ++v) {
          const int i = i0 + v * kTfThreads;
          if (i >= n || !tf_usable(tv[v], xv[v])) continue;
          ++myv;
          myi0 = min(myi0, i);
          if (A.stat) {
            mysx += (double)xv[v];
            mysxx += (double)xv[v] * (double)xv[v];
          }
          const double u = tf_fold<kVar>(tv[v], P, invP);
          atomicAdd(&lcur[min(nb - 1, (int)(u * bscale))], 1);
        }
      }
      atomicAdd(&s_cnt, myv);
      atomicMin(&s_i0, myi0);
    }
    if (A.stat) {
      // the curve's standard deviation over its usable samples, reduced in a
      // fixed order (wave shuffles, then the waves in turn): every item of a
      // curve gets the same bits
      for (int o = kWave / 2; o > 0; o >>= 1) {
        mysx += __shfl_xor(mysx, o);
        mysxx += __shfl_xor(mysxx, o);
      }
      if ((tid & (kWave - 1)) == 0) {
        L.red_key[tid / kWave] = mysx;
        L.red_nb[tid / kWave] = mysxx;
      }
    }
    __syncthreads();
    const int nv = s_cnt;
    if (A.stat && tid == 0) {  // read by thread 0 only, when it writes the item
      double sx = 0.0, sxx = 0.0;
      for (int w = 0; w < kTfThreads / kWave; ++w) { sx += L.red_key[w]; sxx += L.red_nb[w]; }
      const double cn = (double)nv;
      s_sigma = nv > 1 ? sqrt(fmax(sxx - sx * sx / cn, 0.0) / (cn - 1.0)) : NAN;
    }
    // x_ref = flux of the first usable sample (any constant cancels in the statistic)
    const double xref = (s_i0 < n) ? (double)xload(s_i0) : 0.0;
    {  // exclusive scan of lcur[0..nb) in place (per-thread runs + wave scan)
      const int per = (nb + kTfThreads - 1) / kTfThreads;
      const int a0 = tid * per, b0 = min(nb, a0 + per);
      int sum = 0;
      for (int k = a0; k < b0; ++k) sum += lcur[k];
      int inc = sum;
      const int lane = tid & (kWave - 1), wv = tid / kWave;
      for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
      }
      if (lane == kWave - 1) s_wsum[wv] = inc;
      __syncthreads();
      int run = inc - sum;
#pragma unroll
      for (int w = 0; w < kTfThreads / kWave; ++w) run += w < wv ? s_wsum[w] : 0;
      for (int k = a0; k < b0; ++k) { const int v = lcur[k]; lcur[k] = run; run += v; }
      if (tid == 0) lcur[nb] = nv;
      __syncthreads();
    }
    // scatter; afterwards lcur[b] is the end of bucket b
    for (int i0 = tid; i0 < n; i0 += kTfUnroll * kTfThreads) {
      float tv[kTfUnroll], xv[kTfUnroll];
#pragma unroll
      for (int v = 0; v < kTfUnroll; ++v) {
        const int i = min(i0 + v * kTfThreads, n - 1);
        tv[v] = tc[i];
        xv[v] = xload(i);
      }
#pragma unroll
      for (int v = 0; v < kTfUnroll; ++v) {
        if (i0 + v * kTfThreads >= n || !tf_usable(tv[v], xv[v])) continue;
        const double u = tf_fold<kVar>(tv[v], P, invP);
        const int pos = atomicAdd(&lcur[min(nb - 1, (int)(u * bscale))], 1);
        rec[pos] = TfRec{u, xv[v], 0};
      }
    }
    __syncthreads();
    // bucket-aligned sort chunks of at most kTfSortChunk samples (greedy, from
    // the LDS bucket ends), then the ends go to the workspace
    if (tid == 0) {
      int nbrk = 0, bb = 0;
      s_brk[nbrk++] = 0;
      while (bb < nb && nbrk < kTfMaxBreaks) {
        const int s0 = bb == 0 ? 0 : lcur[bb - 1];
        int lo_b = bb + 1, hi_b = min(nb, bb + kTfSortChunk);
        while (lo_b < hi_b) {
          const int mid = (lo_b + hi_b + 1) >> 1;
          if (lcur[mid - 1] - s0 <= kTfSortChunk) lo_b = mid; else hi_b = mid - 1;
        }
        bb = lo_b;
        s_brk[nbrk++] = bb;
      }
      s_nbrk = (bb >= nb) ? nbrk : -1;  // -1: more chunks than the table holds
    }
    for (int b = tid; b <= nb; b += kTfThreads) cursor[b] = lcur[b];
    __syncthreads();
    // sort every bucket by (u, x)
    if (s_nbrk > 0) {
      double *cu = reinterpret_cast<double *>(&L);
      float *cx = reinterpret_cast<float *>(cu + kTfSortChunk);
      int *cb = reinterpret_cast<int *>(cx + kTfSortChunk);  // bucket ends of the chunk
      for (int c = 0; c + 1 < s_nbrk; ++c) {
        const int bb = s_brk[c], be = s_brk[c + 1];
        const int s0 = bb == 0 ? 0 : cursor[bb - 1];
        const int s1 = cursor[be - 1];
        const int cnt = s1 - s0;
        if (cnt <= kTfSortChunk && be - bb <= kTfSortChunk) {
          for (int k0 = tid; k0 < cnt; k0 += kTfUnroll * kTfThreads) {
            double uv[kTfUnroll];
            float xv[kTfUnroll];
#pragma unroll
            for (int v = 0; v < kTfUnroll; ++v) {
              const int k = min(k0 + v * kTfThreads, cnt - 1);
              const TfRec rv = rec[s0 + k];
              uv[v] = rv.u;
              xv[v] = rv.x;
            }
#pragma unroll
            for (int v = 0; v < kTfUnroll; ++v) {
              const int k = k0 + v * kTfThreads;
              if (k < cnt) { cu[k] = uv[v]; cx[k] = xv[v]; }
            }
          }
          for (int k = tid; k < be - bb; k += kTfThreads) cb[k] = cursor[bb + k] - s0;
          __syncthreads();
          // rank of every sample inside its bucket under the total order
          // (u, key(x), position); equal (u, key(x)) pairs are the same bits,
          // so their relative order never shows in the output
          for (int k = tid; k < cnt; k += kTfThreads) {
            const double ku = cu[k];
            const float kx = cx[k];
            const int kk = tf_xkey(kx);
            const int b = min(be - bb - 1, max(0, min(nb - 1, (int)(ku * bscale)) - bb));
            const int a1 = b == 0 ? 0 : cb[b - 1], a2 = cb[b];
            int rank = 0;
            for (int j = a1; j < a2; ++j) {
              const double ju = cu[j];
              const int jk = tf_xkey(cx[j]);
              // bitwise, not short-circuit: no branch per comparison
              rank += (int)((ju < ku) | ((ju == ku) & ((jk < kk) | ((jk == kk) & (j < k)))));
            }
            const int dst = s0 + min(a1 + rank, cnt - 1);
            rec[dst] = TfRec{ku, kx, 0};
          }
          __syncthreads();
        } else {
          // one bucket larger than a chunk (strongly clustered phases; a
          // chunk of several buckets holds at most kTfSortChunk samples)
          tf_bitonic(rec + s0, cnt);
        }
      }
    } else {
      // more chunks than the chunk table holds (a very long curve): one
      // bitonic sort of the whole folded curve orders every bucket too,
      // since the bucket index is non-decreasing in u
      tf_bitonic(rec, nv);
    }

    // ---- 2/3. epoch tiles ---------------------------------------------
    TfBest best{-INFINITY, NAN, NAN, 0ull};
    // first guess of the tile size from the mean phase density
    const double per_epoch = (double)nv / P * A.dtep;
    int te = (int)(((double)kTfChunk / fmax(per_epoch, 1e-9) - 2.0 * (hwm + marg) / A.dtep) - 3.0);
    te = max(1, min(A.max_tile, te));
    int e0 = 0;
    while (e0 < nE) {
      const int e1 = min(nE, e0 + te);
      // union range of the windows of epochs [e0, e1) from the bucket offsets
      // (one bucket of slack each side; the exact boundary searches exclude
      // samples outside a window)
      const double tlo = (double)e0 * A.dtep, thi = (double)(e1 - 1) * A.dtep;
      const int lo = tf_bucket_bound(cursor, nb, nv, P, bscale, tlo - hwm - marg, 0);
      const int hi = tf_bucket_bound(cursor, nb, nv, P, bscale, thi + hwm + marg, 1);
      const int cnt = hi - lo;
      if (cnt > kTfChunk && e1 - e0 > 1) {  // too dense here: halve the tile
        te = max(1, (e1 - e0) / 2);
        continue;
      }
      if (cnt <= kTfChunk) {
        const double r = (double)((e0 + e1 - 1) / 2) * A.dtep;  // q reference: central epoch
        tf_stage<kVar>(L, rec, nv, P, lo, cnt, r, xref, A.cells);
        // the staged tile is swept in runs of te_run epochs
        for (int eb = e0; eb < e1; eb += A.te_run) {
          const int nte = min(A.te_run, e1 - eb);
          // (epoch, W) window table: p0 (d >= -W/2), pE (d >= W/2), pM (d > W/2)
          for (int k = tid; k < nte * nw; k += kTfThreads) {
            const int ep = eb + (k == tid ? wt_e : k / nw), w = k == tid ? wt_w : k % nw;
            const double t0 = (double)ep * A.dtep, W = sw[w];
            // p0: d >= -W/2; pE: d >= W/2 and pM: d > W/2 share one bracket
            int a0, b0, a1, b1;
            tf_bracket(L, t0, -0.5 * W, a0, b0);
            tf_bracket(L, t0, 0.5 * W, a1, b1);
            int p0 = tf_bisect(L, a0, b0, t0, -0.5 * W, 0);
            int pE = tf_bisect(L, a1, b1, t0, 0.5 * W, 0);
            int pM = tf_bisect(L, a1, b1, t0, 0.5 * W, 1);
            if ((b0 - a0 > kTfCount) | (b1 - a1 > kTfCount)) {
              if (b0 - a0 > kTfCount) p0 = tf_search(L.lu, a0, b0, t0, -0.5 * W, 0);
              if (b1 - a1 > kTfCount) {
                pE = tf_search(L.lu, a1, b1, t0, 0.5 * W, 0);
                pM = tf_search(L.lu, a1, b1, t0, 0.5 * W, 1);
              }
            }
            p0 = tf_clamp(p0, 0, cnt);
            pE = tf_clamp(pE, p0, cnt);
            pM = tf_clamp(pM, pE, cnt);
            // W >= P is no trial (the window would hold two images of a
            // sample): an empty window makes every trial of it NaN
            if (!(W < P)) p0 = pE = pM = 0;
            ewp[3 * k] = p0; ewp[3 * k + 1] = pE; ewp[3 * k + 2] = pM;
            tf_window((double)(pM - p0), L.px[pM][0] - L.px[p0][0], (double)(pE - p0),
                      L.px[pE][0] - L.px[p0][0], L.px[pE][1] - L.px[p0][1], ewd[2 * k],
                      ewd[2 * k + 1]);
          }
          __syncthreads();
          // one thread per (epoch, W, L >= 1) trial, L fastest; the thread's
          // (L, W, epoch) indices advance by kTfThreads trials per step
          const int ntr = nte * nw * nlt;
          int mm = nlt ? tid % nlt : 0, w = nlt ? (tid / nlt) % nw : 0,
              je = nlt ? tid / (nlt * nw) : 0;
          for (int k = tid; k < ntr; k += kTfThreads) {
            // row offsets by 24-bit multiplies (all operands < 2^24)
            const unsigned j = __umul24((unsigned)je, (unsigned)nw) + (unsigned)w;
            const int ep = eb + je, m = mm + 1;
            const double t0 = (double)ep * A.dtep, s = t0 - r, W = sw[w];
            const int *wp = ewp + __umul24(j, 3u);
            const double *wd = ewd + 2 * j;
            const int p0 = wp[0], pE = wp[1];
            const double *bd =
                wlt + __umul24((unsigned)kTfWl, __umul24((unsigned)w, (unsigned)nl) + (unsigned)m);
            const double cb[4] = {bd[0], bd[1], bd[2], bd[3]};
            int ba[4], bb[4], br[4];
            // the phase-index coordinate of t0 + c is f0 + c tinv
            const double tinv = L.tinv, f0 = (t0 - L.tbase) * tinv;
#pragma unroll
            for (int i = 0; i < 4; ++i) tf_bracket_f(L, fma(cb[i], tinv, f0), ba[i], bb[i]);
            // bracket widths are >= 0: one of them exceeds 3 (or kTfCount = 7)
            // iff their bitwise or does
            const int wor = (bb[0] - ba[0]) | (bb[1] - ba[1]) | (bb[2] - ba[2]) | (bb[3] - ba[3]);
            tf_count4(L, ba, t0, cb, wor > 3, br);
            if (wor > kTfCount) {
#pragma unroll
              for (int i = 0; i < 4; ++i)
                if (bb[i] - ba[i] > kTfCount) br[i] = tf_search(L.lu, ba[i], bb[i], t0, cb[i], 0);
            }
            const int pAB = tf_clamp(br[0], p0, pE);
            const int pBC = tf_clamp(br[1], pAB, pE);
            const int pCD = tf_clamp(br[2], pBC, pE);
            const int pDE = tf_clamp(br[3], pCD, pE);
            double nB, dA;
            const double key =
                tf_stat_px(L, p0, pAB, pBC, pCD, pDE, pE, wd[0], wd[1], s, W, bd, A.stat, &nB,
                           &dA);
            const unsigned long long idx = tf_idx(w, m, ep, nl);
            if (tf_better(key, idx, best)) best = TfBest{key, nB, dA, idx};
            // next trial (wrap by an unsigned min: m - n wraps above m when m < n)
            mm += dm;
            const int c1 = mm >= nlt ? 1 : 0;
            mm = (int)min((unsigned)mm, (unsigned)(mm - nlt));
            w += dw + c1;
            const int c2 = w >= nw ? 1 : 0;
            w = (int)min((unsigned)w, (unsigned)(w - nw));
            je += dje + c2;
          }
          __syncthreads();
        }
      } else {
        // ---- wide pass: a window larger than the LDS tile (dense items).
        // The union of the windows of up to tf_pass_slots / (nw nbnd) epochs
        // is staged chunk by chunk, the prefix sums carried from chunk to
        // chunk (one reference q = image - r for the whole pass).  Every
        // (epoch, W, boundary) is resolved once, in the chunk holding its
        // first passing sample, and its prefix row goes to the workspace;
        // the trials then read their seven rows from there.
        const int nkind = nw * nbnd;
        // epochs per pass from the curve's own length (n <= nmax: the
        // workspace holds them), so an item's rounding is the same in any batch
        const int tce = max(1, min(nE - e0, tf_pass_slots(n) / nkind));
        const double r = (double)(e0 + (tce - 1) / 2) * A.dtep;  // central epoch
        const int chi = tf_bucket_bound(cursor, nb, nv, P, bscale,
                                        (double)(e0 + tce - 1) * A.dtep + hwm + marg, 1);
        double carry[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double prev = -INFINITY;  // image of the previous chunk's last sample
        for (int ca = lo; ca < chi; ca += kTfChunk) {
          const int cc = min(kTfChunk, chi - ca);
          const bool first = ca == lo, last = ca + cc >= chi;
          tf_stage<kVar>(L, rec, nv, P, ca, cc, r, xref, A.cells);
          const double ptop = L.lu[cc - 1];
          for (int kd = tid; kd < nkind; kd += kTfThreads) {
            const int w = kd / nbnd, b = kd - w * nbnd;
            double c;
            int kind;
            tf_bnd(sw[w], wlt + kTfWl * w * nl, b, c, kind);
            // candidate epochs: boundary value t0 + c in [prev, ptop], with
            // a margin of 1e-6 epoch steps each side (the exact test decides)
            int ea = e0, eb = e0 + tce - 1;
            if (!first) ea = max(ea, (int)ceil((prev - c) / A.dtep - 1e-6));
            if (!last) eb = min(eb, (int)floor((ptop - c) / A.dtep + 1e-6));
            for (int e = ea; e <= eb; ++e) {
              const double t0 = (double)e * A.dtep;
              int a, bb;
              tf_bracket(L, t0, c, a, bb);
              int p = tf_bisect(L, a, bb, t0, c, kind);
              if (bb - a > kTfCount) p = tf_search(L.lu, a, bb, t0, c, kind);
              p = tf_clamp(p, 0, cc);
              // the boundary is in this chunk when its first passing sample
              // is: inside the chunk (or past the last chunk's end), and at
              // the chunk's start only if the previous chunk's last sample
              // fails the test
              bool take = p < cc || last;
              if (p == 0 && !first) {
                const double d = prev - t0;
                take = !(kind ? d > c : d >= c);
              }
              if (take) {
                const double *row = tf_px(L, p);
                TfSlot sl;
#pragma unroll
                for (int v = 0; v < 5; ++v) sl.row[v] = carry[v] + row[v];
                sl.idx = ca - lo + p;
                sl.pad = 0;
                slots[(size_t)(e - e0) * nkind + kd] = sl;
              }
            }
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) carry[v] += L.px[cc][v];
          prev = ptop;
          __syncthreads();
        }
        // one thread per (epoch, W, L >= 1) trial of the pass
        for (int k = tid; k < tce * nw * nlt; k += kTfThreads) {
          const int mm = k % nlt, jj = k / nlt, m = mm + 1;
          const int el = jj / nw, w = jj % nw;
          const double W = sw[w];
          const TfSlot *sb = slots + (size_t)el * nkind + w * nbnd;
          const TfSlot *s0 = sb, *sE = sb + 1, *sM = sb + 2, *sAB = sb + 3 + 4 * mm,
                       *sBC = sAB + 1, *sCD = sAB + 2, *sDE = sAB + 3;
          double nB = NAN, dA = NAN, key = __builtin_nan("");
          if (W < P) {  // W >= P: no trial
            double mean, y2;
            tf_window((double)(sM->idx - s0->idx), sM->row[0] - s0->row[0],
                      (double)(sE->idx - s0->idx), sE->row[0] - s0->row[0],
                      sE->row[1] - s0->row[1], mean, y2);
            const double t0 = (double)(e0 + el) * A.dtep;
            key = tf_stat_rows(s0->idx, sAB->idx, sBC->idx, sCD->idx, sDE->idx, sE->idx, s0->row,
                               sAB->row, sBC->row, sCD->row, sDE->row, sE->row, mean, y2, t0 - r,
                               W, wlt + kTfWl * (w * nl + m), A.stat, &nB, &dA);
          }
          const unsigned long long idx = tf_idx(w, m, e0 + el, nl);
          if (tf_better(key, idx, best)) best = TfBest{key, nB, dA, idx};
        }
        __syncthreads();
        e0 += tce;
        te = 1;
        continue;
      }
      e0 = e1;
      te = min(A.max_tile, 2 * te);
    }
    // ---- reduce the per-thread bests (wave shuffles, then LDS) ----
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const double ok = __shfl_xor(best.key, o), onb = __shfl_xor(best.numB, o),
                   oda = __shfl_xor(best.denA, o);
      const unsigned long long oidx = __shfl_xor(best.idx, o);
      if (oidx != 0ull && tf_better(ok, oidx, best)) best = TfBest{ok, onb, oda, oidx};
    }
    if ((tid & (kWave - 1)) == 0) {
      const int wv = tid / kWave;
      L.red_key[wv] = best.key; L.red_nb[wv] = best.numB; L.red_da[wv] = best.denA;
      L.red_idx[wv] = best.idx;
    }
    __syncthreads();
    if (tid == 0) {
      TfBest b{L.red_key[0], L.red_nb[0], L.red_da[0], L.red_idx[0]};
      for (int k = 1; k < kTfThreads / kWave; ++k)
        if (L.red_idx[k] != 0ull && tf_better(L.red_key[k], L.red_idx[k], b))
          b = TfBest{L.red_key[k], L.red_nb[k], L.red_da[k], L.red_idx[k]};
      const bool found = b.idx != 0ull;
      const int wl = found ? (int)((b.idx - 1ull) >> 32) : -1;
      A.sn[item] = A.stat ? tf_sn(b.key) / s_sigma : tf_sn(b.key);
      if (A.hbest) A.hbest[item] = found ? tf_h(b.numB, b.denA) : NAN;
      A.bw[item] = found ? wl / nl : -1;
      A.bl[item] = found ? wl % nl : -1;
      A.be[item] = found ? (int)((b.idx - 1ull) & 0xffffffffull) : -1;
    }
    __syncthreads();
  }
}

static __global__ void tf_zero_ticket(int *p) { *p = 0; }

// Execution geometry of the launches (gtrap_tlsfold_set_tiles): tile epochs,
// phase-index cells per sample, samples per phase bucket, window-table run
// (0: as large as the LDS allows).
struct TfTiles {
  int max_tile = kTfMaxTile, cells = 2, spb = 2, te_run = 0;
};
extern TfTiles g_tiles;  // defined in tlsfold.hip

// Dynamic LDS of a launch: the (W, L) table and widths, then te_run epochs
// of window table (2 doubles + 3 ints per (epoch, W)), te_run as large as
// the LDS left over by the static part allows (at most kTfMaxTe).
template <bool kItems, int kVar>
static int tf_launch_cfg(int nw, int nl, size_t *dyn, int *te_run) {
  static size_t static_lds = 0;  // per instantiation
  if (!static_lds) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(tlsfold_kernel<kItems, kVar>)) !=
        hipSuccess) {
      set_error("tlsfold: hipFuncGetAttributes failed");
      return GTRAP_ELAUNCH;
    }
    static_lds = fa.sharedSizeBytes;
  }
  const size_t budget = 160 * 1024 > static_lds ? 160 * 1024 - static_lds : 0;
  const size_t fixed = ((size_t)kTfWl * nw * nl + nw) * sizeof(double);
  const size_t per = (size_t)nw * (2 * sizeof(double) + 3 * sizeof(int));
  if (budget < fixed + per) {
    set_error("tlsfold: nw x nl too large for LDS");
    return GTRAP_EUNSUPPORTED;
  }
  int te = (int)((budget - fixed) / per);
  if (te > kTfMaxTe) te = kTfMaxTe;
  if (g_tiles.te_run > 0 && g_tiles.te_run < te) te = g_tiles.te_run;
  *te_run = te;
  *dyn = fixed + (size_t)te * per;
  return GTRAP_OK;
}

}  // namespace gtrap
