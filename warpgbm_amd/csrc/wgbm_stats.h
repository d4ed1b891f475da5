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

// Fold one wave's values into a node's stats (lanes with ok == false skip).
__device__ __forceinline__ void stats_add_wave(int32_t* s, bool ok, float g, float h) {
  int ga = ok ? float_ord(fabsf(g)) : 0;
  int hmn = ok ? float_ord(h) : 0x7fffffff;
  int hmx = ok ? float_ord(h) : (int)0x80000000;
  int ha = ok ? float_ord(fabsf(h)) : 0;
  for (int o = 32; o >= 1; o >>= 1) {
    ga = max(ga, __shfl_xor(ga, o));
    hmn = min(hmn, __shfl_xor(hmn, o));
    hmx = max(hmx, __shfl_xor(hmx, o));
    ha = max(ha, __shfl_xor(ha, o));
  }
  if ((threadIdx.x & 63) == 0 && hmn <= hmx) {  // some lane had a row
    atomicMax(s + 0, ga);
    atomicMin(s + 1, hmn);
    atomicMax(s + 2, hmx);
    atomicMax(s + 3, ha);
  }
}

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

// Bits of the per-row fixed-point gradient and hessian.
constexpr int kGBits = 28;
constexpr int kHBits = 48;
// Packed LDS word of the gradient plane: count << 44 | gradient sum (two's
// complement in the low 44 bits; a chunk's |sum| < 2^(28+14) < 2^43).
constexpr int kCountShift = 44;
constexpr int kMaxChunkRows = 1 << 14;

}  // namespace wgbm
