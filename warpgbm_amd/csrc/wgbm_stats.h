// Per-node gradient statistics and the fixed-point encoding of histograms.
//
// Histograms accumulate integers: a row's gradient becomes
// qg = round(g * 2^sg) and its hessian qh = round(h * 2^sh), with the
// shifts chosen from the node's max |g| and max |h| so that |qg| < 2^34
// (2^49 without a packed count) and |qh| < 2^48.  Integer sums are exact and independent of the order of the
// additions, so any number of waves may add into one LDS histogram and any
// chunking gives the same bits; each bin's sum is rounded to float32 once.
// A node's statistics (max |g|, min h, max h, max |h|) are kept as
// order-preserving integers so they can be merged with atomic min/max.
#pragma once
#include "common.h"

namespace wgbm {

// Stats words per node: {max |g| bits, ord(min h), ord(max h), max |h| bits}.
constexpr int kStatWords = 4;

__host__ __device__ __forceinline__ int float_ord(float x) {
  const int b = __builtin_bit_cast(int, x);
  return b >= 0 ? b : (b ^ 0x7fffffff);
}

__host__ __device__ __forceinline__ float ord_float(int o) {
  return __builtin_bit_cast(float, o >= 0 ? o : (o ^ 0x7fffffff));
}

__device__ __forceinline__ void stats_init(int32_t* s) {
  s[0] = 0;
  s[1] = 0x7fffffff;
  s[2] = (int)0x80000000;
  s[3] = 0;
}

// A thread's running statistics.  Blocks reduce theirs and store one
// partial per block (plain stores: thousands of same-address atomics from
// every wave of a grid serialise at ~10 ns each); k_stats_reduce then folds
// the partials of each node.
struct StatAcc {
  int ga = 0, hmn = 0x7fffffff, hmx = (int)0x80000000, ha = 0;

  __device__ __forceinline__ void add(float g, float h) {
    ga = max(ga, float_ord(fabsf(g)));
    hmn = min(hmn, float_ord(h));
    hmx = max(hmx, float_ord(h));
    ha = max(ha, float_ord(fabsf(h)));
  }
  __device__ __forceinline__ void merge(int a0, int a1, int a2, int a3) {
    ga = max(ga, a0);
    hmn = min(hmn, a1);
    hmx = max(hmx, a2);
    ha = max(ha, a3);
  }
  __device__ __forceinline__ void load_merge(const int32_t* s) { merge(s[0], s[1], s[2], s[3]); }
  __device__ __forceinline__ void store(int32_t* s) const {
    s[0] = ga; s[1] = hmn; s[2] = hmx; s[3] = ha;
  }
  __device__ __forceinline__ void wave_reduce() {
    for (int o = 32; o >= 1; o >>= 1)
      merge(__shfl_xor(ga, o), __shfl_xor(hmn, o), __shfl_xor(hmx, o), __shfl_xor(ha, o));
  }
  // Whole-block reduction (blockDim.x a multiple of 64, at most 1024; every
  // thread calls it).  Thread 0 holds the result.
  __device__ __forceinline__ void block_reduce() {
    __shared__ int32_t sh[16][4];
    wave_reduce();
    const int w = threadIdx.x >> 6;
    __syncthreads();  // sh may still be read by a previous call
    if ((threadIdx.x & 63) == 0) store(sh[w]);
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 1; i < (int)(blockDim.x >> 6); ++i) load_merge(sh[i]);
  }
};

// Shift for values bounded by 2^bits: largest s with max * 2^s < 2^bits.
__device__ __forceinline__ int fixed_shift(float maxabs, int bits) {
  if (!(maxabs > 0.0f) || !(maxabs < INFINITY)) return 0;
  return bits - 1 - ilogbf(maxabs);
}

__device__ __forceinline__ long long to_fixed(float x, int shift) {
  return llrint(ldexp((double)x, shift));
}

__device__ __forceinline__ float from_fixed(long long s, int shift) {
  return (float)ldexp((double)s, -shift);
}

// Gradient words.  A node whose rows share one hessian value ("count
// mode": its H is count * h) packs the row count into the top 15 bits of
// the gradient word (a chunk holds at most 2^14 rows) above a 34-bit
// fixed-point gradient; any other node keeps a count-free 49-bit gradient
// (its hessian has a plane of its own).  Either way a chunk's LDS sum stays
// within 64 bits; chunk partials are summed in 128-bit integers.
constexpr int kCountShift = 49;
constexpr int kGBitsCount = 34;
constexpr int kGBitsWide = 49;
constexpr int kMaxChunkRows = 1 << 14;
// Fine rows (count mode only).  A row whose 34-bit word would carry a
// relative rounding error above 2^-18 -- g != 0 with |round(g * 2^sg)| <
// 2^17 -- adds only its count to the LDS word; its gradient goes, 2^x times
// finer (x = fine_extra), into an exact 64-bit global accumulator per
// (slot, feature, bin) that the read-out adds in.
constexpr long long kFineLimit = 1ll << 17;
// x for a bin matrix of n_rows rows: a fine word is below 2^(18 + x), so a
// bin's fine sum over any segment (at most n_rows rows) stays below 2^63.
__host__ __device__ __forceinline__ int fine_extra(int64_t n_rows) {
  int len = 0;
  while (len < 62 && (int64_t(1) << len) <= n_rows) ++len;  // bits of n_rows
  const int x = 45 - len;
  return x < 30 ? (x > 1 ? x : 1) : 30;
}
// Bits of the per-row fixed-point hessian for a bin matrix of n_rows rows:
// a bin's sum over any segment (at most n_rows rows) stays below 2^62, and a
// chunk's LDS word below 2^62 at the cap of 48.  Matrices of up to 2^14 rows
// keep all 48 bits.
__host__ __device__ __forceinline__ int hess_bits(int64_t n_rows) {
  int len = 0;
  while (len < 62 && (int64_t(1) << len) <= n_rows) ++len;  // bits of n_rows
  const int b = 62 - len;
  return b < 48 ? (b > 16 ? b : 16) : 48;
}

// A signed 128-bit integer as two 64-bit halves (chunk partials of a
// segment are summed exactly in it).
struct I128 {
  unsigned long long lo = 0;
  long long hi = 0;
  __device__ __forceinline__ void add(long long v) {
    const unsigned long long s = lo + (unsigned long long)v;
    hi += (v < 0 ? -1 : 0) + (s < lo ? 1 : 0);
    lo = s;
  }
  __device__ __forceinline__ void add(const I128& o) {
    const unsigned long long s = lo + o.lo;
    hi += o.hi + (s < lo ? 1 : 0);
    lo = s;
  }
  // Round to nearest double.  A value that fits 64 bits converts as that
  // int64 does, so a sum read here and the same sum read from one chunk's
  // 64-bit word give the same double.
  __device__ __forceinline__ double to_double() const {
    if (hi == ((long long)lo >> 63)) return (double)(long long)lo;
    const bool neg = hi < 0;
    unsigned long long h = (unsigned long long)hi, l = lo;
    if (neg) { l = ~l + 1ull; h = ~h + (l == 0ull ? 1ull : 0ull); }
    // |x| >= 2^63: keep its top 63 bits with the rest folded into a sticky
    // low bit (10 bits beyond a double's 53, so one rounding is exact).
    const int p = h ? 127 - __builtin_clzll(h) : 63 - __builtin_clzll(l);  // leading bit, >= 63
    const int sh = p - 62;  // 1 .. 65
    unsigned long long top, rest;
    if (sh < 64) {
      top = (h << (64 - sh)) | (l >> sh);
      rest = l & ((1ull << sh) - 1ull);
    } else {
      top = h >> (sh - 64);
      rest = l | (h & ((1ull << (sh - 64)) - 1ull));
    }
    const double m = (double)(top | (rest ? 1ull : 0ull));
    const double r = ldexp(m, sh);
    return neg ? -r : r;
  }
};

// Two-stage statistics: at most kStatBlocks per-block partials per node
// (stats_blocks(rows) of them), folded by k_stats_reduce (one block per node;
// part is [node][n_part][kStatWords]).
constexpr int kStatBlocks = 1024;
int stats_blocks(int64_t n);
__global__ void k_gh_stats(const float2* __restrict__ gh, int64_t n, int32_t* __restrict__ part);
__global__ void k_list_stats(const int32_t* __restrict__ rows, int n, int64_t n_rows, const float* __restrict__ g,
                             const float* __restrict__ h, int32_t* __restrict__ part);
__global__ void k_stats_reduce(const int32_t* __restrict__ part, int n_part, int32_t* __restrict__ out);

}  // namespace wgbm
