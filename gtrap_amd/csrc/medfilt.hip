// Median detrend (gfilter) and the on-device offset (get_offset) for gfx950.
//
// Reference semantics: gtrap/gfilter.py:13-85.  For each output row i the
// reference forms the window x~[i-r .. i+r] with zero padding outside
// 0 <= i < n (FILVAL 0.0, gfilter.py:12,25-36), selects its exact median by
// value bisection (gfilter.py:40-74) and writes x[i] - median (gfilter.py:76).
// Any exact selection returns the same window element, so this file uses a
// different selector: every lane owns one (curve, time segment) and keeps the
// window sorted in registers while it slides along time.  One slide step
// deletes the outgoing sample and inserts the incoming one with 3 branch-free
// VALU ops per slot (compare, select, v_med3_f32); the median is the centre
// slot.  Windows narrower than the register array are padded with -inf/+inf
// sentinels so the median always sits at compile-time index RMAX.
//
// Lanes are numbered g = segment*cols + curve, so for the interleaved (n, nq)
// image consecutive lanes read consecutive curves: one wave reads 256
// contiguous bytes per row (coalesced), unlike the reference's (1,1000)
// blocks that read one curve per block with stride nq.
//
// The batch workload's radii r = 24 and r = 64 run a blocked form instead:
// the window rows shared by B consecutive outputs are ordered once by a
// sorting network and each output's median is selected from them and its
// B-1 own rows (fewer VALU ops per output than sliding the whole window).
//
// Windows wider than 129 slots (r > 64, the reference drivers' r = 100, 128
// and 512) keep one sorted window across a group of 2-16 lanes (129 slots
// each, DPP moves between neighbours); windows beyond 16 lanes (r > 1031) use
// an exact bisection over the float key space on LDS tiles.
#include "common.h"
#include <math.h>
#include <utility>

namespace gtrap {

// Delete `ov` (known to be present) from the sorted array s and insert `nv`.
// d[k] = s[k] < ov ? s[k] : s[k+1] is the array with ov deleted (+inf at the
// end); the new s[k] = med3(d[k-1], nv, d[k]) (s[0] = min(nv, d[0])) inserts
// nv.  Slot k is rewritten as soon as d[k] and d[k+1] exist, so only two d
// values are live at a time and the window needs about W + 4 registers (the
// whole d array kept 2W + a few live: one wave per SIMD at W = 129).
template <int W>
__device__ __forceinline__ void slide_sorted(float (&s)[W], float ov, float nv) {
  static_assert(W > 1, "a window of at least two slots");
  float dprev = (s[0] < ov) ? s[0] : s[1];
  float snew = fminf(nv, dprev);  // the new s[0]
#pragma unroll
  for (int k = 1; k < W; ++k) {
    const float nxt = (k + 1 < W) ? s[k + 1 < W ? k + 1 : k] : INFINITY;
    const float dcur = (s[k] < ov) ? s[k] : nxt;
    s[k - 1] = snew;  // old s[k-1] was last read by d[k-1]
    snew = med3(dprev, nv, dcur);
    dprev = dcur;
  }
  s[W - 1] = snew;
}

constexpr int kUnroll = 8;

// Filter rows [i0, i1) of one curve whose sample i sits at xq[i*stride]
// (n samples, zero padding outside [0, n)); the sorted window lives in
// registers and slides one row at a time.
template <int RMAX>
__device__ __forceinline__ void medfilt_run(const float *__restrict__ xq, float *__restrict__ yq,
                                            size_t stride, int n, int i0, int i1, int r,
                                            int mode) {
  constexpr int W = 2 * RMAX + 1;
  float s[W];
#pragma unroll
  for (int k = 0; k < W; ++k)
    s[k] = (k < RMAX - r) ? -INFINITY : ((k > RMAX + r) ? INFINITY : 0.0f);
  auto ld = [&](int i) -> float { return (i >= 0 && i < n) ? xq[(size_t)i * stride] : 0.0f; };

  // The 2r+1 middle slots start as placeholder zeros; replace them one by one
  // with the window of row i0 (loads issued kUnroll at a time).
  for (int j = -r; j <= r; j += kUnroll) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = (j + u <= r) ? ld(i0 + j + u) : 0.0f;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (j + u <= r) slide_sorted<W>(s, 0.0f, v[u]);
  }

  // Main sweep, software-pipelined: the loads of the next kUnroll rows are in
  // flight while the current rows are filtered (the sweep is otherwise bound
  // by one dependent global-load latency per row).
  float cin[kUnroll], cout_[kUnroll], cc[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    cc[u] = ld(i0 + u);
    cin[u] = ld(i0 + u + r + 1);
    cout_[u] = ld(i0 + u - r);
  }
  for (int i = i0; i < i1; i += kUnroll) {
    float nin[kUnroll], nout[kUnroll], nc[kUnroll];
    const int in_ = i + kUnroll;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool more = in_ < i1;
      nc[u] = more ? ld(in_ + u) : 0.0f;
      nin[u] = more ? ld(in_ + u + r + 1) : 0.0f;
      nout[u] = more ? ld(in_ + u - r) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (i + u < i1) {
        const float med = s[RMAX];
        yq[(size_t)(i + u) * stride] = mode ? cc[u] / med : cc[u] - med;
        slide_sorted<W>(s, cout_[u], cin[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      cc[u] = nc[u];
      cin[u] = nin[u];
      cout_[u] = nout[u];
    }
  }
}

template <int RMAX>
__global__ __launch_bounds__(256) void medfilt_win_kernel(
    const float *__restrict__ x, float *__restrict__ y, int n, int nq, int r, int mode,
    int rows_out, int cols_out, int seg, int nseg) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nseg * cols_out) return;
  const int q = (int)(g % cols_out);
  const int sidx = (int)(g / cols_out);
  const int i0 = sidx * seg;
  const int i1 = min(i0 + seg, rows_out);
  if (i0 >= i1) return;
  medfilt_run<RMAX>(x + q, y + q, (size_t)nq, n, i0, i1, r, mode);
}

// Ragged batch of contiguous curves (serving): lane g filters segment g,
// rows [seg_start[g], seg_start[g]+seg) of curve seg_curve[g]; the segment
// count is read from device memory so the launch shape is fixed.
template <int RMAX>
__global__ __launch_bounds__(256) void medfilt_ragged_kernel(
    const float *__restrict__ x, float *__restrict__ y, const int64_t *__restrict__ off,
    const int32_t *__restrict__ len, const int32_t *__restrict__ seg_curve,
    const int32_t *__restrict__ seg_start, const int32_t *__restrict__ nseg_dev, int r, int mode,
    int seg) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *nseg_dev) return;
  const int c = seg_curve[g];
  const int n = len[c];
  const int i0 = seg_start[g];
  const int i1 = min(i0 + seg, n);
  if (i0 >= i1) return;
  medfilt_run<RMAX>(x + off[c], y + off[c], 1, n, i0, i1, r, mode);
}

// ---- blocked windows (r = 24 and r = 64): B outputs per sorted core ----
//
// The windows (W = 2r + 1 rows) of the B consecutive outputs i .. i+B-1 share
// M = W - (B-1) rows, i+B-1-r .. i+r, the core C.  Output j adds B-1 extra
// rows, E_j = rows i+j-r .. i+B-2-r and i+r+1 .. i+r+j.  With C and E_j each
// sorted, the median (the (r+1)-th smallest of C u E_j) is
//   min over a = 0..B-1 of max(E_j[a-1], C[r-a])     (E_j[-1] = -inf):
// every term is at least the (r+1)-th smallest (a elements of E_j and r+1-a
// of C lie at or below it) and the term with a = the number of E_j elements
// among the r+1 smallest equals it.  Only C[r-B+1 .. r] enter, so per block
// the core is ordered by a sorting network pruned to those ranks, E_0 by a
// full one, and E_{j+1} follows from E_j by the one-lane slide (delete row
// i+j-r, insert row i+r+1+j).  Per output that is 2(B-1) min/max for the
// selection, 3(B-1) for the slide and the networks' share: at r = 64, B = 24
// (1030 + 127 compare-exchanges per block) about 220 VALU ops against the
// slide's 3 W = 387 over the whole window; at r = 24, B = 16 (232 + 59)
// about 115 against 147.

// Compare-exchange list of Batcher's odd-even merge sort over the next power
// of two >= m, less the exchanges that are no-ops or dead: padding slots
// >= m hold +inf, which every exchange (a, b), a < b, leaves in place, so
// exchanges with b >= m are dropped; exchanges whose outputs reach none of
// the ranks lo..hi (backwards from the end) are dropped too.
struct MfNet {
  int n;
  unsigned char a[1600], b[1600];
};
constexpr MfNet mf_make_net(int m, int lo, int hi) {
  MfNet all{0, {}, {}};
  int np2 = 1;
  while (np2 < m) np2 <<= 1;
  for (int p = 1; p < np2; p <<= 1)
    for (int k = p; k >= 1; k >>= 1)
      for (int j = k % p; j + k < np2; j += 2 * k)
        for (int i = 0; i < k && i < np2 - j - k; ++i)
          if ((i + j) / (2 * p) == (i + j + k) / (2 * p) && i + j + k < m) {
            all.a[all.n] = (unsigned char)(i + j);
            all.b[all.n] = (unsigned char)(i + j + k);
            ++all.n;
          }
  bool live[256] = {};
  for (int k = lo; k <= hi; ++k) live[k] = true;
  bool keep[1600] = {};
  int nk = 0;
  for (int c = all.n - 1; c >= 0; --c)
    if (live[all.a[c]] || live[all.b[c]]) {
      keep[c] = true;
      live[all.a[c]] = live[all.b[c]] = true;
      ++nk;
    }
  MfNet net{0, {}, {}};
  for (int c = 0; c < all.n; ++c)
    if (keep[c]) {
      net.a[net.n] = all.a[c];
      net.b[net.n] = all.b[c];
      ++net.n;
    }
  return net;
}
template <int M, int LO, int HI>
struct MfNetT {
  static constexpr MfNet net = mf_make_net(M, LO, HI);
  static_assert(net.n > 0 && net.n <= 1600, "sorting network table");
};

// min and max as v_med3_f32 against -inf/+inf: the same values as fminf/fmaxf
// for non-NaN operands, without the canonicalising v_max_f32 that IEEE-mode
// fminf/fmaxf put in front of operands read from memory.  The infinities come
// from scalar moves the compiler cannot see through (it would fold med3
// against a constant infinity back into fminf/fmaxf).
struct MfInf {
  float lo, hi;
};
__device__ __forceinline__ MfInf mf_inf() {
  MfInf f;
  asm("s_mov_b32 %0, 0xff800000" : "=s"(f.lo));
  asm("s_mov_b32 %0, 0x7f800000" : "=s"(f.hi));
  return f;
}
__device__ __forceinline__ float mf_min(float x, float y, MfInf f) { return med3(x, y, f.lo); }
__device__ __forceinline__ float mf_max(float x, float y, MfInf f) { return med3(x, y, f.hi); }

template <int A, int B, int N>
__device__ __forceinline__ void mf_ce(float (&v)[N], MfInf f) {
  const float x = v[A], y = v[B];
  v[A] = mf_min(x, y, f);
  v[B] = mf_max(x, y, f);
}
template <class T, int N, size_t... C>
__device__ __forceinline__ void mf_apply(float (&v)[N], MfInf f, std::index_sequence<C...>) {
  (mf_ce<T::net.a[C], T::net.b[C]>(v, f), ...);
}

// A float 0.0 in device memory: rows outside the curve read it (the zero
// padding) through the same load as every other row, so edge blocks issue
// their loads together instead of branching around them.
__device__ float mf_zero_row = 0.0f;

// Rows [i0, i1) of one curve (sample i at xq[i*stride]), r = RM (no
// sentinels).  Blocks whose rows i-RM .. i+RM+B all lie inside the curve read
// through running pointers (one 64-bit add per row); blocks at the curve's
// ends read row by row through a bounds check that points outside rows at
// mf_zero_row.
template <int RM, int B>
__device__ __forceinline__ void medfilt_block_run(const float *__restrict__ xq,
                                                  float *__restrict__ yq, size_t stride, int n,
                                                  int i0, int i1, int mode) {
  constexpr int W = 2 * RM + 1, M = W - (B - 1), NE = B - 1;
  static_assert(B % kUnroll == 0, "whole load groups per block");
  using NetC = MfNetT<M, RM - NE, RM>;
  using NetE = MfNetT<NE, 0, NE - 1>;
  const float *zero = &mf_zero_row;
  const MfInf inf = mf_inf();
  // row k's address, or the zero row outside the curve
  auto at = [&](int k) -> const float * {
    return (unsigned)k < (unsigned)n ? xq + (size_t)(unsigned)k * stride : zero;
  };
  for (int i = i0; i < i1; i += B) {
    const bool inner = i - RM >= 0 && i + RM + B < n;  // wave-uniform for a dense image
    // rows i+jg .. +kUnroll-1 (outputs), RM before them (deleted from E) and
    // RM+1 after them (inserted)
    auto load_group = [&](int jg, float (&xc)[kUnroll], float (&xo)[kUnroll],
                          float (&xn)[kUnroll]) {
      if (inner) {
        const float *pc = xq + (size_t)(i + jg) * stride, *po = pc - (size_t)RM * stride,
                    *pn = pc + (size_t)(RM + 1) * stride;
#pragma unroll
        for (int u = 0; u < kUnroll; ++u, pc += stride, po += stride, pn += stride) {
          xc[u] = *pc;
          xo[u] = *po;
          xn[u] = *pn;
        }
      } else {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int j = i + jg + u;
          xc[u] = *at(j);
          xo[u] = *at(j - RM);
          xn[u] = *at(j + RM + 1);
        }
      }
    };
    // core slot s holds row i+B-1-RM+s; E_0 rows i-RM .. i+B-2-RM
    float c[M], e[NE];
    if (inner) {
      const float *p = xq + (size_t)(i - RM) * stride;
#pragma unroll
      for (int t = 0; t < NE; ++t, p += stride) e[t] = *p;
#pragma unroll
      for (int s = 0; s < M; ++s, p += stride) c[s] = *p;
    } else {
#pragma unroll
      for (int t = 0; t < NE; ++t) e[t] = *at(i - RM + t);
#pragma unroll
      for (int s = 0; s < M; ++s) c[s] = *at(i + B - 1 - RM + s);
    }
    mf_apply<NetC>(c, inf, std::make_index_sequence<NetC::net.n>{});
    float cm[B];  // C[RM-B+1 .. RM]
#pragma unroll
    for (int t = 0; t < B; ++t) cm[t] = c[RM - NE + t];
    mf_apply<NetE>(e, inf, std::make_index_sequence<NetE::net.n>{});
    // outputs in groups of kUnroll, the group's rows loaded before its first use
    float *py = yq + (size_t)i * stride;
    for (int jg = 0; jg < B; jg += kUnroll) {
      float xc[kUnroll], xo[kUnroll], xn[kUnroll];
      load_group(jg, xc, xo, xn);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u, py += stride) {
        float med = cm[NE];
#pragma unroll
        for (int k = 1; k <= NE; ++k) med = mf_min(med, mf_max(e[k - 1], cm[NE - k], inf), inf);
        if (i + jg + u < i1) *py = mode ? xc[u] / med : xc[u] - med;
        slide_sorted<NE>(e, xo[u], xn[u]);  // E_{j+1} (unused after the block's last output)
      }
    }
  }
}

// dense image, r = RM: lane g = segment*cols + curve filters rows
// [segment*seg, +seg) (seg a multiple of B)
template <int RM, int B>
__global__ __launch_bounds__(256) void medfilt_block_kernel(
    const float *__restrict__ x, float *__restrict__ y, int n, int nq, int mode, int rows_out,
    int cols_out, int seg, int nseg) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nseg * cols_out) return;
  const int q = (int)(g % cols_out);
  const int sidx = (int)(g / cols_out);
  const int i0 = sidx * seg;
  const int i1 = min(i0 + seg, rows_out);
  if (i0 >= i1) return;
  medfilt_block_run<RM, B>(x + q, y + q, (size_t)nq, n, i0, i1, mode);
}

// ragged batch (serving), r = RM: lane g filters segment g of the segment list
template <int RM, int B>
__global__ __launch_bounds__(256) void medfilt_ragged_block_kernel(
    const float *__restrict__ x, float *__restrict__ y, const int64_t *__restrict__ off,
    const int32_t *__restrict__ len, const int32_t *__restrict__ seg_curve,
    const int32_t *__restrict__ seg_start, const int32_t *__restrict__ nseg_dev, int mode,
    int seg) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *nseg_dev) return;
  const int c = seg_curve[g];
  const int n = len[c];
  const int i0 = seg_start[g];
  const int i1 = min(i0 + seg, n);
  if (i0 >= i1) return;
  medfilt_block_run<RM, B>(x + off[c], y + off[c], 1, n, i0, i1, mode);
}

// ---- wide windows in registers: one sorted window over a group of G lanes ----
//
// Lane gp (0..G-1) of a group holds slots [gp*WL, (gp+1)*WL) of one sorted
// array of T = G*WL slots; the median sits at slot T/2 - 1, the last slot of
// lane G/2 - 1, so windows up to r = T/2 - 1 fit (the rest are -inf/+inf
// sentinels, as in the one-lane window).  A slide is the one-lane slide over
// the whole array: slot k becomes med3(d[k-1], nv, d[k]) with d[k] = s[k] if
// s[k] < ov else s[k+1]; a lane needs only its right neighbour's first slot
// (for its last d) and its left neighbour's last d (for its first slot), two
// DPP moves inside a row of 16 lanes (G divides 16, groups never straddle a
// row).  Work per output is 3 T VALU ops spread over G lanes.

// lane i reads lane i+1 (row_shl:1) or lane i-1 (row_shr:1) of its row
__device__ __forceinline__ float dpp_from_next(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_from_prev(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}

template <int WL, int G>
__device__ __forceinline__ void slide_group(float (&s)[WL], float ov, float nv, int gp) {
  float nxt0 = dpp_from_next(s[0]);
  if (gp == G - 1) nxt0 = INFINITY;
  const float dlast = (s[WL - 1] < ov) ? s[WL - 1] : nxt0;
  float dprev = dpp_from_prev(dlast);
  if (gp == 0) dprev = -INFINITY;  // med3(-inf, nv, d[0]) = min(nv, d[0])
#pragma unroll
  for (int k = 0; k < WL; ++k) {
    const float dcur = (k + 1 < WL) ? ((s[k] < ov) ? s[k] : s[k + 1 < WL ? k + 1 : k]) : dlast;
    s[k] = med3(dprev, nv, dcur);  // old s[k] was last read by dcur
    dprev = dcur;
  }
}

// Rows [i0, i1) of one curve (sample i at xq[i*stride]) by the group of G
// lanes this lane (position gp) belongs to; every lane of the group runs the
// same rows, and lane G/2 - 1 writes the outputs.
template <int WL, int G>
__device__ __forceinline__ void medfilt_group_run(const float *__restrict__ xq,
                                                  float *__restrict__ yq, size_t stride, int n,
                                                  int i0, int i1, int r, int mode, int gp) {
  constexpr int T = G * WL, RM = T / 2 - 1;
  float s[WL];
#pragma unroll
  for (int k = 0; k < WL; ++k) {
    const int gk = gp * WL + k;
    s[k] = (gk < RM - r) ? -INFINITY : ((gk > RM + r) ? INFINITY : 0.0f);
  }
  auto ld = [&](int i) -> float { return (i >= 0 && i < n) ? xq[(size_t)i * stride] : 0.0f; };
  for (int j = -r; j <= r; j += kUnroll) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = (j + u <= r) ? ld(i0 + j + u) : 0.0f;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (j + u <= r) slide_group<WL, G>(s, 0.0f, v[u], gp);
  }
  const bool writer = gp == G / 2 - 1;
  float cin[kUnroll], cout_[kUnroll], cc[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    cc[u] = ld(i0 + u);
    cin[u] = ld(i0 + u + r + 1);
    cout_[u] = ld(i0 + u - r);
  }
  for (int i = i0; i < i1; i += kUnroll) {
    float nin[kUnroll], nout[kUnroll], nc[kUnroll];
    const int in_ = i + kUnroll;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool more = in_ < i1;
      nc[u] = more ? ld(in_ + u) : 0.0f;
      nin[u] = more ? ld(in_ + u + r + 1) : 0.0f;
      nout[u] = more ? ld(in_ + u - r) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (i + u < i1) {
        if (writer) {
          const float med = s[WL - 1];
          yq[(size_t)(i + u) * stride] = mode ? cc[u] / med : cc[u] - med;
        }
        slide_group<WL, G>(s, cout_[u], cin[u], gp);
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      cc[u] = nc[u];
      cin[u] = nin[u];
      cout_[u] = nout[u];
    }
  }
}

// dense image: group g = segment*cols + curve, lanes g*G .. g*G + G-1
template <int WL, int G>
__global__ __launch_bounds__(256) void medfilt_group_kernel(
    const float *__restrict__ x, float *__restrict__ y, int n, int nq, int r, int mode,
    int rows_out, int cols_out, int seg, int nseg) {
  const long lane = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long g = lane / G;
  const int gp = (int)(lane % G);
  if (g >= (long)nseg * cols_out) return;  // whole groups (blockDim is a multiple of G)
  const int q = (int)(g % cols_out);
  const int sidx = (int)(g / cols_out);
  const int i0 = sidx * seg;
  const int i1 = min(i0 + seg, rows_out);
  if (i0 >= i1) return;
  medfilt_group_run<WL, G>(x + q, y + q, (size_t)nq, n, i0, i1, r, mode, gp);
}

// ragged batch: group g filters segment g of the segment list
template <int WL, int G>
__global__ __launch_bounds__(256) void medfilt_ragged_group_kernel(
    const float *__restrict__ x, float *__restrict__ y, const int64_t *__restrict__ off,
    const int32_t *__restrict__ len, const int32_t *__restrict__ seg_curve,
    const int32_t *__restrict__ seg_start, const int32_t *__restrict__ nseg_dev, int r, int mode,
    int seg) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = lane / G, gp = lane % G;
  if (g >= *nseg_dev) return;
  const int c = seg_curve[g];
  const int n = len[c];
  const int i0 = seg_start[g];
  const int i1 = min(i0 + seg, n);
  if (i0 >= i1) return;
  medfilt_group_run<WL, G>(x + off[c], y + off[c], 1, n, i0, i1, r, mode, gp);
}

// Order-preserving map float -> uint32 (total order, -0 < +0).
__device__ __forceinline__ uint32_t fkey(float v) {
  uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}

// Windows beyond the lane groups (r > 1031): exact (r+1)-th smallest by
// bisection over the 32-bit key space, 32 counting passes over the window per
// output.
__global__ __launch_bounds__(256) void medfilt_wide_kernel(
    const float *__restrict__ x, float *__restrict__ y, int n, int nq, int r, int mode,
    int rows_out, int cols_out) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)rows_out * cols_out) return;
  const int q = (int)(g % cols_out);
  const int i = (int)(g / cols_out);
  const int lo = i - r, hi = i + r;
  uint32_t key = 0;  // smallest key K with count(key <= K) >= r+1
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t trial = key | ((1u << bit) - 1u);  // all lower bits set
    int cnt = 0;
    for (int j = lo; j <= hi; ++j) {
      const float v = (j >= 0 && j < n) ? x[(size_t)j * nq + q] : 0.0f;
      cnt += (fkey(v) <= trial) ? 1 : 0;
    }
    if (cnt < r + 1) key |= (1u << bit);
  }
  const float med = fkey_inv(key);
  const float xv = x[(size_t)i * nq + q];
  y[(size_t)i * nq + q] = mode ? xv / med : xv - med;
}

// Wide windows in tiles: a workgroup filters rows [i0, i1) (at most
// kWideTile) of one curve whose rows are x[i*stride] (i < n).  It stages
// rows i0 - r .. i1 - 1 + r once into LDS as order-preserving keys (zero
// outside [0, n), the reference's halo), then every thread finds its
// window's (r+1)-th smallest key by a 32-step bisection over the key bits,
// counting from LDS (consecutive threads read consecutive words).  The
// result is the exact window median, as every selector returns the same
// element.
constexpr int kWideTile = 256;
constexpr int kWideMaxR = 4096;  // LDS: (kWideTile + 2r) keys

__device__ void medfilt_wide_tile(const float *__restrict__ x, float *__restrict__ y,
                                  long stride, int n, int i0, int i1, int r, int mode,
                                  uint32_t *keys) {
  const int span = (i1 - i0) + 2 * r;
  for (int k = threadIdx.x; k < span; k += blockDim.x) {
    const int j = i0 - r + k;
    keys[k] = fkey((j >= 0 && j < n) ? x[(long)j * stride] : 0.0f);
  }
  __syncthreads();
  const int i = i0 + (int)threadIdx.x;
  if (i < i1) {
    const uint32_t *wk = keys + threadIdx.x;  // window of row i: wk[0 .. 2r]
    uint32_t key = 0;  // smallest K with count(key <= K) >= r + 1
    for (int bit = 31; bit >= 0; --bit) {
      const uint32_t trial = key | ((1u << bit) - 1u);
      int cnt = 0;
      for (int j = 0; j <= 2 * r; ++j) cnt += (wk[j] <= trial) ? 1 : 0;
      if (cnt < r + 1) key |= (1u << bit);
    }
    const float med = fkey_inv(key);
    const float xv = x[(long)i * stride];
    y[(long)i * stride] = mode ? xv / med : xv - med;
  }
  __syncthreads();
}

// dense image (interleaved (n, nq)): grid (ceil(rows_out / kWideTile), cols_out)
__global__ __launch_bounds__(kWideTile) void medfilt_wide_tile_kernel(
    const float *__restrict__ x, float *__restrict__ y, int n, int nq, int r, int mode,
    int rows_out) {
  extern __shared__ uint32_t wkeys[];
  const int q = blockIdx.y;
  const int i0 = blockIdx.x * kWideTile;
  const int i1 = min(rows_out, i0 + kWideTile);
  medfilt_wide_tile(x + q, y + q, nq, n, i0, i1, r, mode, wkeys);
}

// ragged segments: block b filters segment b (curve slot seg_curve[b], rows
// seg_start[b] .. + seg) in tiles of kWideTile rows
__global__ __launch_bounds__(kWideTile) void medfilt_ragged_wide_kernel(
    const float *__restrict__ x, float *__restrict__ y, const int64_t *__restrict__ off,
    const int32_t *__restrict__ len, const int32_t *__restrict__ seg_curve,
    const int32_t *__restrict__ seg_start, const int32_t *__restrict__ nseg_dev, int r, int mode,
    int seg) {
  extern __shared__ uint32_t wkeys[];
  const int b = blockIdx.x;
  if (b >= *nseg_dev) return;
  const int c = seg_curve[b];
  const int n = len[c];
  const int s0 = seg_start[b];
  const int s1 = min(s0 + seg, n);
  for (int i0 = s0; i0 < s1; i0 += kWideTile)
    medfilt_wide_tile(x + off[c], y + off[c], 1, n, i0, min(s1, i0 + kWideTile), r, mode,
                      wkeys);
}

// Per-curve masked mean over n (gtrap/gfilter.py:94-101): one block per curve,
// float64 accumulation, fixed reduction tree (deterministic).
__global__ __launch_bounds__(256) void offset_kernel(const float *__restrict__ img,
                                                     const float *__restrict__ tu, int n,
                                                     int nq, double *off64, float *off32,
                                                     uint32_t *ntrue) {
  const int q = blockIdx.x;
  double acc = 0.0;
  uint32_t cnt = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const size_t e = (size_t)i * nq + q;
    if (tu[e] >= 0.0f) {
      acc += (double)img[e];
      cnt += 1;
    }
  }
  __shared__ double sacc[256];
  __shared__ uint32_t scnt[256];
  sacc[threadIdx.x] = acc;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sacc[threadIdx.x] += sacc[threadIdx.x + s];
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = sacc[0] / (double)n;
    if (off64) off64[q] = m;
    if (off32) off32[q] = (float)m;
    if (ntrue) ntrue[q] = scnt[0];
  }
}

template <int RMAX>
static void launch_win(const float *in, float *out, int n, int nq, int r, int mode,
                       int rows_out, int cols_out, int seg, int nseg, hipStream_t st) {
  const long lanes = (long)nseg * cols_out;
  const int block = 256;
  const long grid = (lanes + block - 1) / block;
  hipLaunchKernelGGL((medfilt_win_kernel<RMAX>), dim3((unsigned)grid), dim3(block), 0, st, in,
                     out, n, nq, r, mode, rows_out, cols_out, seg, nseg);
}

// Windows wider than the one-lane buckets (r > 64): groups of G lanes of
// kGroupWL slots each, the smallest G in {2, 4, 8, 16} with r <= G*kGroupWL/2 - 1;
// wider windows take the LDS bisection.
constexpr int kGroupWL = 129;
constexpr int kGroupMaxR = 16 * kGroupWL / 2 - 1;
static int group_lanes(int r) {
  int G = 2;
  while (G < 16 && r > G * kGroupWL / 2 - 1) G <<= 1;
  return G;
}

// Register-window buckets: the smallest RMAX >= r is used.  r = 24 and r = 64
// themselves (the detrend radii of the batch workload) run the blocked
// selection with kBlockLanes lanes: 1.5 rounds of waves at r = 64 (156 VGPRs,
// three waves per SIMD), about one at r = 24 (96 VGPRs, five); measured
// within 3% of the best of 196608 .. 524288 lanes for both.
constexpr long kBlockLanes = 393216;
static const int kBuckets[] = {4, 8, 16, 24, 32, 48, 64};

int medfilt_launch(const float *in, float *out, uint32_t n, uint32_t nq, uint32_t r, int mode,
                   uint32_t rows_out, uint32_t cols_out, uint32_t seg_len, hipStream_t st) {
  GTRAP_REQUIRE(in && out, "gtrap_medfilt: null pointer");
  GTRAP_REQUIRE(mode == GTRAP_MEDFILT_SUB || mode == GTRAP_MEDFILT_DIV,
                "gtrap_medfilt: mode must be 0 (sub) or 1 (div)");
  GTRAP_REQUIRE(rows_out <= n && cols_out <= nq, "gtrap_medfilt: rows_out/cols_out exceed n/nq");
  GTRAP_REQUIRE(n < (1u << 30) && nq < (1u << 30) && r < (1u << 28),
                "gtrap_medfilt: size out of range");
  GTRAP_REQUIRE(in != out, "gtrap_medfilt: in-place filtering is not supported");
  if (rows_out == 0 || cols_out == 0) return GTRAP_OK;
  const int w = 2 * (int)r + 1;
  if (r > 64 && r <= (uint32_t)kGroupMaxR) {
    const int G = group_lanes((int)r);
    int seg = (int)seg_len;
    if (seg <= 0) {
      // 131072 lanes (eight waves per CU, the two per SIMD that 193 VGPRs
      // allow: one full round), at least 4 windows per segment
      const long target = 131072 / G;
      long per = (long)cols_out >= target ? 1 : (target + cols_out - 1) / cols_out;
      seg = (int)(((long)rows_out + per - 1) / per);
      if (seg < 4 * w) seg = 4 * w;
    }
    if (seg > (int)rows_out) seg = (int)rows_out;
    const int nseg = ((int)rows_out + seg - 1) / seg;
    const long lanes = (long)nseg * cols_out * G;
    const dim3 grid((unsigned)((lanes + 255) / 256));
#define GTRAP_MF_GROUP(GG)                                                                      \
  hipLaunchKernelGGL((medfilt_group_kernel<kGroupWL, GG>), grid, dim3(256), 0, st, in, out,      \
                     (int)n, (int)nq, (int)r, mode, (int)rows_out, (int)cols_out, seg, nseg)
    switch (G) {
      case 2: GTRAP_MF_GROUP(2); break;
      case 4: GTRAP_MF_GROUP(4); break;
      case 8: GTRAP_MF_GROUP(8); break;
      default: GTRAP_MF_GROUP(16); break;
    }
#undef GTRAP_MF_GROUP
    return check_launch("medfilt_group_kernel");
  }
  if (r > 64 && r <= (uint32_t)kWideMaxR && cols_out <= 65535) {
    const size_t shm = (size_t)(kWideTile + 2 * r) * sizeof(uint32_t);
    dim3 grid((rows_out + kWideTile - 1) / kWideTile, cols_out);
    hipLaunchKernelGGL(medfilt_wide_tile_kernel, grid, dim3(kWideTile), shm, st, in, out, (int)n,
                       (int)nq, (int)r, mode, (int)rows_out);
    return check_launch("medfilt_wide_tile_kernel");
  }
  if (r > 64) {
    const long total = (long)rows_out * cols_out;
    hipLaunchKernelGGL(medfilt_wide_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       st, in, out, (int)n, (int)nq, (int)r, mode, (int)rows_out,
                       (int)cols_out);
    return check_launch("medfilt_wide_kernel");
  }
  if ((int)r == 64 || (int)r == 24) {
    // blocked selection (B = 24 outputs per core at r = 64, 16 at r = 24):
    // kBlockLanes lanes, segments a whole number of blocks
    const int B = (int)r == 64 ? 24 : 16;
    int seg = (int)seg_len;
    if (seg <= 0) {
      const long per = (long)cols_out >= kBlockLanes ? 1 : (kBlockLanes + cols_out - 1) / cols_out;
      seg = (int)(((long)rows_out + per - 1) / per);
      seg = (seg + B - 1) / B * B;
    }
    if (seg > (int)rows_out) seg = (int)rows_out;
    const int nseg = ((int)rows_out + seg - 1) / seg;
    const long lanes = (long)nseg * cols_out;
    const dim3 grid((unsigned)((lanes + 255) / 256));
    if ((int)r == 64)
      hipLaunchKernelGGL((medfilt_block_kernel<64, 24>), grid, dim3(256), 0, st, in, out, (int)n,
                         (int)nq, mode, (int)rows_out, (int)cols_out, seg, nseg);
    else
      hipLaunchKernelGGL((medfilt_block_kernel<24, 16>), grid, dim3(256), 0, st, in, out, (int)n,
                         (int)nq, mode, (int)rows_out, (int)cols_out, seg, nseg);
    return check_launch("medfilt_block_kernel");
  }
  int seg = (int)seg_len;
  if (seg <= 0) {
    // One round of waves (256 CUs x 4 SIMDs x the waves per SIMD that the
    // register window allows: 4 up to RMAX = 32, 2 above), at least 4 windows
    // per segment so the initial fill is amortised.
    int rmax_ = 64;
    for (int b : kBuckets)
      if ((int)r <= b) { rmax_ = b; break; }
    const long target = rmax_ >= 48 ? 131072 : 262144;
    long per = (long)cols_out >= target ? 1 : (target + cols_out - 1) / cols_out;
    seg = (int)(((long)rows_out + per - 1) / per);
    if (seg < 4 * w) seg = 4 * w;
  }
  if (seg > (int)rows_out) seg = (int)rows_out;
  const int nseg = ((int)rows_out + seg - 1) / seg;
  int rmax = 64;
  for (int b : kBuckets)
    if ((int)r <= b) { rmax = b; break; }
  switch (rmax) {
    case 4: launch_win<4>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    case 8: launch_win<8>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    case 16: launch_win<16>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    case 24: launch_win<24>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    case 32: launch_win<32>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    case 48: launch_win<48>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
    default: launch_win<64>(in, out, n, nq, r, mode, rows_out, cols_out, seg, nseg, st); break;
  }
  return check_launch("medfilt_win_kernel");
}

}  // namespace gtrap

extern "C" int gtrap_medfilt(const float *in, float *out, uint32_t n, uint32_t nq, uint32_t r,
                             int mode, uint32_t rows_out, uint32_t cols_out, uint32_t seg_len,
                             void *stream) {
  return gtrap::medfilt_launch(in, out, n, nq, r, mode, rows_out, cols_out, seg_len,
                               gtrap::as_stream(stream));
}

namespace gtrap {

template <int RMAX>
static void launch_ragged(const float *x, float *y, const int64_t *off, const int32_t *len,
                          const int32_t *sc, const int32_t *ss, const int32_t *nseg, int max_segs,
                          int r, int mode, int seg, hipStream_t st) {
  hipLaunchKernelGGL((medfilt_ragged_kernel<RMAX>), dim3((max_segs + 255) / 256), dim3(256), 0,
                     st, x, y, off, len, sc, ss, nseg, r, mode, seg);
}

int medfilt_ragged_launch(const float *x, float *y, const int64_t *off, const int32_t *len,
                          const int32_t *sc, const int32_t *ss, const int32_t *nseg,
                          int32_t max_segs, uint32_t r, int mode, int32_t seg,
                          hipStream_t st) {
  GTRAP_REQUIRE(x && y && off && len && sc && ss && nseg, "gtrap_medfilt_ragged: null pointer");
  GTRAP_REQUIRE(mode == 0 || mode == 1, "gtrap_medfilt_ragged: mode must be 0 or 1");
  GTRAP_REQUIRE(r <= (uint32_t)kWideMaxR, "gtrap_medfilt_ragged: r must be <= %d", kWideMaxR);
  GTRAP_REQUIRE(seg >= 1 && max_segs >= 0, "gtrap_medfilt_ragged: bad segment shape");
  if (max_segs == 0) return GTRAP_OK;
  if (r > 64 && r <= (uint32_t)kGroupMaxR) {
    const int G = group_lanes((int)r);
    const dim3 grid((unsigned)(((long)max_segs * G + 255) / 256));
#define GTRAP_MF_RGROUP(GG)                                                                      \
  hipLaunchKernelGGL((medfilt_ragged_group_kernel<kGroupWL, GG>), grid, dim3(256), 0, st, x, y,   \
                     off, len, sc, ss, nseg, (int)r, mode, seg)
    switch (G) {
      case 2: GTRAP_MF_RGROUP(2); break;
      case 4: GTRAP_MF_RGROUP(4); break;
      case 8: GTRAP_MF_RGROUP(8); break;
      default: GTRAP_MF_RGROUP(16); break;
    }
#undef GTRAP_MF_RGROUP
    return check_launch("medfilt_ragged_group_kernel");
  }
  if (r > 64) {
    const size_t shm = (size_t)(kWideTile + 2 * r) * sizeof(uint32_t);
    hipLaunchKernelGGL(medfilt_ragged_wide_kernel, dim3((unsigned)max_segs), dim3(kWideTile), shm,
                       st, x, y, off, len, sc, ss, nseg, (int)r, mode, seg);
    return check_launch("medfilt_ragged_wide_kernel");
  }
  if ((int)r == 64 || (int)r == 24) {  // blocked selection, as in medfilt_launch
    const dim3 grid((unsigned)((max_segs + 255) / 256));
    if ((int)r == 64)
      hipLaunchKernelGGL((medfilt_ragged_block_kernel<64, 24>), grid, dim3(256), 0, st, x, y, off,
                         len, sc, ss, nseg, mode, seg);
    else
      hipLaunchKernelGGL((medfilt_ragged_block_kernel<24, 16>), grid, dim3(256), 0, st, x, y, off,
                         len, sc, ss, nseg, mode, seg);
    return check_launch("medfilt_ragged_block_kernel");
  }
  int rmax = 64;
  for (int b : kBuckets)
    if ((int)r <= b) { rmax = b; break; }
  switch (rmax) {
    case 4: launch_ragged<4>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    case 8: launch_ragged<8>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    case 16: launch_ragged<16>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    case 24: launch_ragged<24>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    case 32: launch_ragged<32>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    case 48: launch_ragged<48>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
    default: launch_ragged<64>(x, y, off, len, sc, ss, nseg, max_segs, r, mode, seg, st); break;
  }
  return check_launch("medfilt_ragged_kernel");
}

}  // namespace gtrap

extern "C" int gtrap_medfilt_ragged(const float *x, float *y, const int64_t *off,
                                    const int32_t *len, const int32_t *seg_curve,
                                    const int32_t *seg_start, const int32_t *nseg_dev,
                                    int32_t max_segs, uint32_t r, int mode, int32_t seg_len,
                                    void *stream) {
  return gtrap::medfilt_ragged_launch(x, y, off, len, seg_curve, seg_start, nseg_dev, max_segs,
                                      r, mode, seg_len, gtrap::as_stream(stream));
}

extern "C" int gtrap_offset(const float *img, const float *tu, uint32_t n, uint32_t nq,
                            double *offset64, float *offset32, uint32_t *ntrue, void *stream) {
  GTRAP_REQUIRE(img && tu, "gtrap_offset: null pointer");
  GTRAP_REQUIRE(n > 0 && nq > 0, "gtrap_offset: empty image");
  hipLaunchKernelGGL(gtrap::offset_kernel, dim3(nq), dim3(256), 0, gtrap::as_stream(stream), img,
                     tu, (int)n, (int)nq, offset64, offset32, ntrue);
  return gtrap::check_launch("offset_kernel");
}
