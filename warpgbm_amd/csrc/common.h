// Shared helpers for the gfx950 kernels and their C-ABI launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

namespace wgbm {

constexpr int kWave = 64;

// Last error message of the calling host thread (returned by wgbm_last_error).
void set_error(const char* fmt, ...);

// Return codes of every C-ABI entry point.
enum Status : int { kOk = 0, kBadArg = 1, kHipError = 2, kTooBig = 3 };

#define WGBM_REQUIRE(cond, ...)                 \
  do {                                          \
    if (!(cond)) {                              \
      ::wgbm::set_error(__VA_ARGS__);           \
      return ::wgbm::kBadArg;                   \
    }                                           \
  } while (0)

#define WGBM_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::wgbm::set_error("%s failed: %s", #expr, hipGetErrorString(_e));       \
      return ::wgbm::kHipError;                                               \
    }                                                                         \
  } while (0)

#define WGBM_LAUNCHED(name)                                                   \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::wgbm::set_error("launch of %s failed: %s", name, hipGetErrorString(_e)); \
      return ::wgbm::kHipError;                                               \
    }                                                                         \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Largest i in [0, n) with a[i] <= key, for a non-decreasing array a with
// a[0] <= key.
__device__ __forceinline__ int upper_index(const int32_t* __restrict__ a, int n, int key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Population count of the lanes below the calling lane in a 64-bit mask.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Exclusive prefix sum of v over the block (blockDim.x a multiple of 64, at
// most 1024; every thread calls it); *total receives the block's sum.
__device__ __forceinline__ int block_exscan(int v, int* total) {
  __shared__ int ws[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  __syncthreads();  // ws may still be read by a previous call
  if (lane == 63) ws[w] = x;
  __syncthreads();
  int pre = 0, tot = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
    const int c = ws[k];
    if (k < w) pre += c;
    tot += c;
  }
  *total = tot;
  return pre + x - v;
}

}  // namespace wgbm
