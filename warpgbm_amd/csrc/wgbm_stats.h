// Per-node gradient statistics and the fixed-point encoding of histograms.
//
// Histograms accumulate integers: a row's gradient becomes
// qg = round(g * 2^sg) and its hessian qh = round(h * 2^sh), with the
// shifts chosen from the node's max |g| and max |h| so that |qg| < 2^28 and
// |qh| < 2^48.  Integer sums are exact and independent of the order of the
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

// Bits of the per-row fixed-point gradient: a chunk's sum stays below 2^43
// (packed beside the count) and any segment's below 2^63.
constexpr int kGBits = 28;
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
// Packed LDS word of the gradient plane: count << 44 | gradient sum (two's
// complement in the low 44 bits; a chunk's |sum| < 2^(28+14) < 2^43).
constexpr int kCountShift = 44;
constexpr int kMaxChunkRows = 1 << 14;

// Two-stage statistics: at most kStatBlocks per-block partials per node
// (stats_blocks(rows) of them), folded by k_stats_reduce (one block per node;
// part is [node][n_part][kStatWords]).
constexpr int kStatBlocks = 1024;
int stats_blocks(int64_t n);
__global__ void k_gh_stats(const float2* __restrict__ gh, int64_t n, int32_t* __restrict__ part);
__global__ void k_stats_reduce(const int32_t* __restrict__ part, int n_part, int32_t* __restrict__ out);

}  // namespace wgbm
