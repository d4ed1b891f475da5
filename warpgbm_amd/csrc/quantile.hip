// Quantile bin edges without sorting (gfx950).
//
// torch.quantile(col, q) with linear interpolation needs, per edge j, the
// column's order statistics at ranks lo_j = floor(r_j) and hi_j = ceil(r_j),
// r_j = f32(q_j * (n-1)) (warpgbm/core.py:1271-1278).  Instead of sorting
// each column, every requested order statistic is found by a radix select on
// order-preserving 32-bit keys, 8 bits per round (4 rounds): a round
// histograms the next digit of the keys that share each target's prefix so
// far, and each target descends into the bucket that holds its rank.  Keys
// order -0.0 before +0.0, as a radix sort on the bits would.  A column with
// a NaN gets NaN edges (torch propagates NaN through quantile).
//
// The matrix is first transposed to column-major so that one workgroup reads
// its column contiguously; a workgroup per column then runs the rounds with
// its histograms in LDS.
#include "common.h"

namespace wgbm {

constexpr int kSelThreads = 1024;
constexpr int kSelTargets = 64;  // order statistics per pass (histograms [64][256] in LDS)

__device__ __forceinline__ uint32_t order_key(float x) {
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float key_value(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __builtin_bit_cast(float, b);
}

// Interior quantile level j (1..B-1) as torch.linspace(0, 1, B+1) forms it.
__device__ __forceinline__ float quantile_level(int j, int num_bins) {
  const int steps = num_bins + 1;
  const float step = 1.0f / (float)(steps - 1);
  const int half = steps / 2;
  return (j < half) ? fmaf(step, (float)j, 0.0f) : fmaf(-step, (float)(steps - 1 - j), 1.0f);
}

// xt[c][r] = X[r * ld + c] for c < F, through 32 x 32 LDS tiles.
__global__ void __launch_bounds__(256)
k_transpose_f32(const float* __restrict__ X, int64_t n, int F, int64_t ld, float* __restrict__ xt) {
  __shared__ float tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + k;
    const int c = c0 + tx;
    tile[k][tx] = (r < n && c < F) ? X[r * ld + c] : 0.0f;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k;
    const int64_t r = r0 + tx;
    if (r < n && c < F) xt[(int64_t)c * n + r] = tile[tx][k];
  }
}

__global__ void __launch_bounds__(kSelThreads)
k_quantile_select(const float* __restrict__ xt, int64_t n, int num_bins, float* __restrict__ edges,
                  int edge_row_stride) {
  extern __shared__ uint32_t s_hist[];  // [kSelTargets][256]
  __shared__ int64_t s_rank[2 * 127];   // order statistic wanted by target t (sorted ascending)
  __shared__ float s_val[2 * 127];      // its value once found
  __shared__ uint32_t s_prefix[kSelTargets];
  __shared__ int64_t s_rem[kSelTargets];
  __shared__ uint32_t s_dpre[kSelTargets];
  __shared__ int s_grp[kSelTargets];
  __shared__ int s_nd, s_nan;
  const int c = blockIdx.x;
  const float* col = xt + (int64_t)c * n;
  const int ne = num_bins - 1;
  const int T = 2 * ne;
  const int tid = threadIdx.x;
  // targets: (lo_j, hi_j) for every edge, then sorted by rank (counting sort)
  int64_t my = 0;
  if (tid < T) {
    const int j = tid / 2 + 1;
    const float rank = quantile_level(j, num_bins) * (float)(n - 1);
    my = (tid & 1) ? (int64_t)ceilf(rank) : (int64_t)floorf(rank);
    my = my < n - 1 ? my : n - 1;  // (float)(n-1) rounds up above 2^24 rows
  }
  __shared__ int64_t s_tmp[2 * 127];
  if (tid < T) s_tmp[tid] = my;
  if (tid == 0) s_nan = 0;
  __syncthreads();
  if (tid < T) {
    int pos = 0;
    for (int u = 0; u < T; ++u) pos += (s_tmp[u] < my) || (s_tmp[u] == my && u < tid);
    s_rank[pos] = my;
  }
  __syncthreads();
  for (int g0 = 0; g0 < T; g0 += kSelTargets) {
    const int tg = min(kSelTargets, T - g0);
    if (tid < tg) { s_prefix[tid] = 0; s_rem[tid] = s_rank[g0 + tid]; s_grp[tid] = 0; }
    if (tid == 0) { s_nd = 1; s_dpre[0] = 0; }
    __syncthreads();
    for (int round = 0; round < 4; ++round) {
      const int shift = 24 - 8 * round;
      const int nd = s_nd;
      for (int q = tid; q < nd * 256; q += kSelThreads) s_hist[q] = 0;
      __syncthreads();
      int nan_here = 0;
      for (int64_t i = tid; i < n; i += kSelThreads) {
        const float x = col[i];
        if (x != x) { ++nan_here; continue; }
        const uint32_t k = order_key(x);
        int d = 0;
        if (round > 0) {  // which distinct prefix (sorted) does the key carry?
          const uint32_t p = k >> (shift + 8);
          int lo = 0, hi = nd - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_dpre[mid] <= p) lo = mid; else hi = mid - 1;
          }
          d = (s_dpre[lo] == p) ? lo : -1;
        }
        if (d >= 0) atomicAdd(&s_hist[d * 256 + ((k >> shift) & 255u)], 1u);
      }
      if (round == 0 && g0 == 0 && nan_here) atomicAdd(&s_nan, nan_here);
      __syncthreads();
      if (s_nan) break;  // uniform: the column holds a NaN
      if (tid < tg) {  // descend into the bucket holding the target's rank
        const uint32_t* h = s_hist + s_grp[tid] * 256;
        int64_t rem = s_rem[tid], cum = 0;
        int v = 0;
        for (; v < 255; ++v) {
          if (rem < cum + (int64_t)h[v]) break;
          cum += h[v];
        }
        s_prefix[tid] = (s_prefix[tid] << 8) | (uint32_t)v;
        s_rem[tid] = rem - cum;
      }
      __syncthreads();
      if (tid == 0) {  // distinct prefixes, non-decreasing since the ranks are sorted
        int nd2 = 0;
        for (int t = 0; t < tg; ++t) {
          if (nd2 == 0 || s_dpre[nd2 - 1] != s_prefix[t]) s_dpre[nd2++] = s_prefix[t];
          s_grp[t] = nd2 - 1;
        }
        s_nd = nd2;
      }
      __syncthreads();
    }
    if (s_nan) break;
    if (tid < tg) s_val[g0 + tid] = key_value(s_prefix[tid]);
    __syncthreads();
  }
  __syncthreads();
  if (tid < ne) {
    float e;
    if (s_nan) {
      e = __builtin_nanf("");
    } else {
      const int j = tid + 1;
      const float rank = quantile_level(j, num_bins) * (float)(n - 1);
      const float lo_f = floorf(rank), hi_f = ceilf(rank);
      const float w = rank - lo_f;
      // values at ranks lo and hi: find them among the sorted targets
      const int64_t lo = min((int64_t)lo_f, n - 1), hi = min((int64_t)hi_f, n - 1);
      int ia = 0, ib = 0;
      for (int t = 0; t < T; ++t) {
        if (s_rank[t] == lo) ia = t;
        if (s_rank[t] == hi) ib = t;
      }
      const float a = s_val[ia], b = s_val[ib];
      const float d = b - a;
      e = (w < 0.5f) ? fmaf(w, d, a) : fmaf(-d, 1.0f - w, b);
    }
    edges[(int64_t)c * edge_row_stride + tid] = e;
  }
}

}  // namespace wgbm

using namespace wgbm;

extern "C" size_t wgbm_quantile_edges_workspace_bytes(int64_t n_rows, int32_t n_cols) {
  if (n_rows < 0 || n_cols < 0) return 0;
  return (size_t)n_rows * n_cols * sizeof(float) + 256;
}

// Edges [n_cols][edge_row_stride] (first num_bins-1 used) of the columns of a
// row-major float32 matrix X (row stride ld), as torch.quantile computes them.
extern "C" int wgbm_quantile_edges(const float* X, int64_t n_rows, int32_t n_cols, int64_t ld, int32_t num_bins,
                                   float* edges, int32_t edge_row_stride, void* workspace, size_t workspace_bytes,
                                   hipStream_t stream) {
  WGBM_REQUIRE(n_rows >= 1 && n_cols >= 0 && ld >= n_cols, "bad matrix shape");
  WGBM_REQUIRE(num_bins >= 2 && num_bins <= 128 && edge_row_stride >= num_bins - 1, "bad bin count %d", num_bins);
  WGBM_REQUIRE(workspace_bytes >= wgbm_quantile_edges_workspace_bytes(n_rows, n_cols), "quantile workspace too small");
  WGBM_REQUIRE(ceil_div(n_rows, 32) < 0x7fffffff, "too many rows");
  if (n_cols == 0) return kOk;
  float* xt = static_cast<float*>(workspace);
  hipLaunchKernelGGL(k_transpose_f32, dim3((unsigned)ceil_div(n_rows, 32), (unsigned)ceil_div(n_cols, 32)), dim3(256),
                     0, stream, X, n_rows, (int)n_cols, ld, xt);
  WGBM_LAUNCHED("k_transpose_f32");
  const size_t lds = (size_t)kSelTargets * 256 * sizeof(uint32_t);
  static bool opted[64] = {};
  int dev = 0;
  WGBM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64 || !opted[dev]) {
    WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_quantile_select),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (dev >= 0 && dev < 64) opted[dev] = true;
  }
  hipLaunchKernelGGL(k_quantile_select, dim3((unsigned)n_cols), dim3(kSelThreads), lds, stream, xt, n_rows,
                     (int)num_bins, edges, (int)edge_row_stride);
  WGBM_LAUNCHED("k_quantile_select");
  return kOk;
}
