// Post-search peak statistics of S/N series (one series per curve).
//
// Reference semantics: the statistics block of gtrap/gtls.py:244-275 and
// kepler/gtls_mockkepler.py:336-376 over x = tlssn[k*nq + q], k < nt, with
// the peak finder gtrap/detect_peaks.py:10-138 (edge 'rising', minimum peak
// distance mpd):
//   * candidates: 1 <= i <= nt-2, x[i] > x[i-1], x[i+1] <= x[i], no NaN in
//     x[i-1..i+1];
//   * candidates in descending height (ties: larger index first) are kept
//     greedily, each kept peak removing every candidate within mpd samples;
//   * maxsn = max(x) (NaN if any NaN), median(x) (NaN if any NaN), std =
//     population std of x[x > 0], far = (maxsn - median)/std, lent =
//     count(x > 0), npeak, the top ntop peaks in descending height,
//     diff = (x[peak[ntop-1]] - median)/std and pssn = (x[peak[0]] -
//     median(tail))/std(tail), tail = the peaks after the top ntop (when
//     npeak > ntop; else 0 and +inf), pinterval = |t0[peak[1]] - t0[peak[0]]|.
//
// MI355X design: one workgroup per curve (persistent over curves).  Sums in
// float64 with block reductions; the median by a 4-pass radix select on
// order-preserving 32-bit keys with an LDS histogram; candidates appended
// as 64-bit priority keys (height key, index) to the workgroup's workspace
// and bitonic-sorted (in LDS when they fit, else in the workspace); the
// greedy suppression walks the sorted keys with one wave (64 keys per load,
// broadcast by lane shuffles) against an LDS bitmap of covered samples.
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kPkThreads = 512;
constexpr int kPkLdsKeys = 8192;      // bitonic sort in LDS up to this many keys
constexpr int kPkMaxNt = 131072;      // bitmap of covered samples: nt bits
constexpr int kPkMaxTop = 16;

__device__ __forceinline__ uint32_t pk_key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float pk_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct PkArgs {
  const float *sn, *t0;
  int32_t nt, nq, mpd, ntop;
  double *st;       // [nq][8]: maxsn, median, std, far, diff, pssn, pinterval, unused
  int32_t *cnt;     // [nq][2]: npeak, lent
  int32_t *top;     // [nq][ntop]
  uint64_t *ws;     // per workgroup: kpad keys
  int64_t kpad;
};

template <typename T>
__device__ T pk_block_sum(T v, T *red) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  T s = 0;
  for (int w = 0; w < kPkThreads / kWave; ++w) s += red[w];
  return s;
}

// ascending bitonic sort of n (a power of two) keys
__device__ void pk_bitonic(uint64_t *a, int64_t n) {
  for (int64_t k = 2; k <= n; k <<= 1) {
    for (int64_t j = k >> 1; j > 0; j >>= 1) {
      for (int64_t i = threadIdx.x; i < n; i += kPkThreads) {
        const int64_t l = i ^ j;
        if (l > i) {
          const uint64_t x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[l] = x; }
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kPkThreads) void peakstats_kernel(PkArgs A) {
  __shared__ double red_d[kPkThreads / kWave];
  __shared__ int red_i[kPkThreads / kWave];
  __shared__ float red_f[kPkThreads / kWave];
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_sel, s_rank;
  __shared__ int s_k;
  __shared__ uint64_t lkeys[kPkLdsKeys];  // also the covered-sample bitmap (nt bits)
  const int tid = threadIdx.x;
  const int nt = A.nt, nq = A.nq;
  uint64_t *keys = A.ws + (size_t)blockIdx.x * (size_t)A.kpad;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    auto X = [&](int k) { return A.sn[(size_t)k * nq + q]; };
    // ---- maximum, NaN, positive count and mean --------------------------
    float mx = -INFINITY;
    int anynan = 0, npos = 0;
    double spos = 0.0;
    for (int k = tid; k < nt; k += kPkThreads) {
      const float v = X(k);
      if (!(v == v)) anynan = 1;
      mx = fmaxf(mx, v);
      if (v > 0.0f) { ++npos; spos += (double)v; }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) red_f[tid / kWave] = mx;
    __syncthreads();
    mx = red_f[0];
    for (int w = 1; w < kPkThreads / kWave; ++w) mx = fmaxf(mx, red_f[w]);
    anynan = pk_block_sum<int>(anynan, red_i);
    npos = pk_block_sum<int>(npos, red_i);
    spos = pk_block_sum<double>(spos, red_d);
    const double mean = npos ? spos / npos : 0.0;
    double sq = 0.0;
    for (int k = tid; k < nt; k += kPkThreads) {
      const float v = X(k);
      if (v > 0.0f) { const double d = (double)v - mean; sq += d * d; }
    }
    sq = pk_block_sum<double>(sq, red_d);
    const double stdv = npos ? sqrt(sq / npos) : __longlong_as_double(0x7ff8000000000000LL);
    const double maxsn = anynan ? __longlong_as_double(0x7ff8000000000000LL) : (double)mx;
    // ---- median: radix select of ranks (nt-1)/2 and nt/2 ------------------
    double median = __longlong_as_double(0x7ff8000000000000LL);
    if (!anynan && nt > 0) {
      float sel[2];
      for (int s = 0; s < 2; ++s) {
        const uint32_t rank = s == 0 ? (uint32_t)((nt - 1) / 2) : (uint32_t)(nt / 2);
        uint32_t prefix = 0u, mask = 0u;
        if (tid == 0) s_rank = rank;
        for (int shift = 24; shift >= 0; shift -= 8) {
          for (int b = tid; b < 256; b += kPkThreads) hist[b] = 0u;
          __syncthreads();
          for (int k = tid; k < nt; k += kPkThreads) {
            const uint32_t key = pk_key(X(k));
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
          }
          __syncthreads();
          if (tid == 0) {
            uint32_t r = s_rank, b = 0;
            while (b < 255 && hist[b] <= r) { r -= hist[b]; ++b; }
            s_rank = r;
            s_sel = b;
          }
          __syncthreads();
          prefix |= s_sel << shift;
          mask |= 255u << shift;
          __syncthreads();
        }
        sel[s] = pk_unkey(prefix);
      }
      median = (nt & 1) ? (double)sel[0] : ((double)sel[0] + (double)sel[1]) / 2.0;
    }
    // ---- candidates as ascending priority keys ----------------------------
    if (tid == 0) s_k = 0;
    __syncthreads();
    const bool lds = A.kpad <= kPkLdsKeys;
    uint64_t *kk = lds ? lkeys : keys;
    for (int i = 1 + tid; i < nt - 1; i += kPkThreads) {
      const float a = X(i - 1), b = X(i), c = X(i + 1);
      if (a == a && b == b && c == c && b > a && c <= b) {
        const int slot = atomicAdd(&s_k, 1);
        kk[slot] = ((uint64_t)(0xffffffffu - pk_key(b)) << 32) | (uint64_t)(0xffffffffu - (uint32_t)i);
      }
    }
    __syncthreads();
    const int K = s_k;
    int64_t npad = 1;
    while (npad < K) npad <<= 1;
    for (int64_t i = K + tid; i < npad; i += kPkThreads) kk[i] = ~0ull;
    __syncthreads();
    if (K > 1) pk_bitonic(kk, npad);
    if (lds) {  // sorted keys to the workspace; the LDS becomes the bitmap
      for (int i = tid; i < K; i += kPkThreads) keys[i] = lkeys[i];
      __syncthreads();
    }
    // ---- greedy suppression (one wave), survivors in priority order -------
    uint32_t *cov = reinterpret_cast<uint32_t *>(lkeys);
    const int nw32 = (nt + 31) / 32;
    for (int w = tid; w < nw32; w += kPkThreads) cov[w] = 0u;
    __syncthreads();
    if (tid < kWave) {
      volatile uint32_t *vc = cov;
      int S = 0;
      const int mpd = A.mpd > 1 ? A.mpd : 0;
      for (int j0 = 0; j0 < K; j0 += kWave) {
        const int pos = j0 + tid < K ? (int)(0xffffffffu - (uint32_t)keys[j0 + tid]) : -1;
        const int m = min(kWave, K - j0);
        for (int l = 0; l < m; ++l) {
          const int p = __shfl(pos, l);
          if ((vc[p >> 5] >> (p & 31)) & 1u) continue;
          if (tid == 0) keys[S] = (uint64_t)p;
          ++S;
          if (mpd > 0) {
            const int lo = max(0, p - mpd), hi = min(nt - 1, p + mpd);
            if (tid == 0)
              for (int w = lo >> 5; w <= (hi >> 5); ++w) {
                const int b0 = max(lo, w << 5) - (w << 5), b1 = min(hi, (w << 5) + 31) - (w << 5);
                const uint32_t bits = (b1 == 31 ? 0xffffffffu : ((1u << (b1 + 1)) - 1u)) &
                                      ~((1u << b0) - 1u);
                vc[w] = vc[w] | bits;
              }
          }
        }
      }
      if (tid == 0) s_k = S;
    }
    __syncthreads();
    const int S = s_k;
    // ---- statistics over the peaks ----------------------------------------
    const int ntop = A.ntop;
    if (tid < ntop) A.top[(size_t)q * ntop + tid] = tid < S ? (int32_t)keys[tid] : -1;
    double diff = 0.0, pssn = INFINITY;
    if (S > ntop) {
      const int m = S - ntop;
      double s1 = 0.0;
      for (int i = tid; i < m; i += kPkThreads) s1 += (double)X((int)keys[ntop + i]);
      s1 = pk_block_sum<double>(s1, red_d);
      const double tm = s1 / m;
      double s2 = 0.0;
      for (int i = tid; i < m; i += kPkThreads) {
        const double d = (double)X((int)keys[ntop + i]) - tm;
        s2 += d * d;
      }
      s2 = pk_block_sum<double>(s2, red_d);
      // the tail is in descending height: its median is the middle element(s)
      const double medc = (m & 1) ? (double)X((int)keys[ntop + m / 2])
                                  : ((double)X((int)keys[ntop + m / 2 - 1]) +
                                     (double)X((int)keys[ntop + m / 2])) / 2.0;
      pssn = ((double)X((int)keys[0]) - medc) / sqrt(s2 / m);
      diff = ((double)X((int)keys[ntop - 1]) - median) / stdv;
    }
    if (tid == 0) {
      double *o = A.st + (size_t)q * 8;
      o[0] = maxsn;
      o[1] = median;
      o[2] = stdv;
      o[3] = (maxsn - median) / stdv;
      o[4] = diff;
      o[5] = pssn;
      o[6] = (A.t0 && S >= 2)
                 ? fabs((double)A.t0[(size_t)keys[1] * nq + q] - (double)A.t0[(size_t)keys[0] * nq + q])
                 : __longlong_as_double(0x7ff8000000000000LL);
      o[7] = 0.0;
      A.cnt[(size_t)q * 2] = S;
      A.cnt[(size_t)q * 2 + 1] = npos;
    }
    __syncthreads();
  }
}

}  // namespace gtrap

extern "C" int64_t gtrap_peakstats_ws_bytes(int32_t nt, int32_t nblocks) {
  int64_t kpad = 1;
  while (kpad < (int64_t)nt / 2 + 1) kpad <<= 1;
  return kpad * 8 * (int64_t)(nblocks > 0 ? nblocks : 1);
}

extern "C" int gtrap_peakstats(double *st, int32_t *cnt, int32_t *top, const float *sn,
                               const float *t0, int32_t nt, int32_t nq, int32_t mpd,
                               int32_t ntop, void *ws, int64_t ws_bytes, int32_t nblocks,
                               void *stream) {
  using namespace gtrap;
  GTRAP_REQUIRE(st && cnt && top && sn && ws, "gtrap_peakstats: null pointer");
  GTRAP_REQUIRE(nt >= 1 && nt <= kPkMaxNt && nq >= 1,
                "gtrap_peakstats: need 1 <= nt <= %d, nq >= 1", kPkMaxNt);
  GTRAP_REQUIRE(ntop >= 1 && ntop <= kPkMaxTop, "gtrap_peakstats: ntop must be 1..%d",
                kPkMaxTop);
  GTRAP_REQUIRE(mpd >= 1 && nblocks >= 1, "gtrap_peakstats: need mpd >= 1, nblocks >= 1");
  GTRAP_REQUIRE(ws_bytes >= gtrap_peakstats_ws_bytes(nt, nblocks),
                "gtrap_peakstats: workspace too small");
  static_assert(kPkMaxNt / 8 <= kPkLdsKeys * 8, "bitmap exceeds the LDS key area");
  PkArgs A{sn, t0, nt, nq, mpd, ntop, st, cnt, top, reinterpret_cast<uint64_t *>(ws),
           gtrap_peakstats_ws_bytes(nt, 1) / 8};
  const int g = nblocks < nq ? nblocks : nq;
  hipLaunchKernelGGL(peakstats_kernel, dim3((unsigned)g), dim3(kPkThreads), 0,
                     as_stream(stream), A);
  return check_launch("peakstats_kernel");
}
