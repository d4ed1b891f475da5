// Feature binning to int8 codes (gfx950).
//
// code(x) = number of leading edges e_k with x >= e_k, stopping at the first
// edge greater than x; NaN compares false and yields 0
// (warpgbm/cuda/binner.cu:16-25).  Two entry points: the reference's
// per-column op, and a whole-matrix op that bins an [N, F] row-major float32
// matrix against edges [F, B-1] in one launch, reading 16 bytes of X and
// writing 4 codes per thread (no per-feature launches, no transposes).
#include "common.h"
#include <algorithm>

namespace wgbm {

__device__ __forceinline__ int8_t bin_one(float x, const float* __restrict__ e, int nb, int stride) {
  int b = 0;
  while (b < nb && x >= e[(int64_t)b * stride]) ++b;
  return (int8_t)b;
}

__global__ void __launch_bounds__(256)
k_bin_column(const float* __restrict__ x, const float* __restrict__ edges, int8_t* __restrict__ out,
             int64_t n, int nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = bin_one(x[i], edges, nb, 1);
}

// Edges transposed to [B-1][F] so that threads of consecutive features read
// consecutive edge words.
__global__ void __launch_bounds__(256)
k_transpose_edges(const float* __restrict__ edges, float* __restrict__ et, int F, int nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)F * nb) return;
  const int f = (int)(i / nb), k = (int)(i - (int64_t)f * nb);
  et[(int64_t)k * F + f] = edges[i];
}

// Four consecutive elements of a row per thread (F % 4 == 0).
__global__ void __launch_bounds__(256)
k_bin_matrix4(const float4* __restrict__ X, const float* __restrict__ et, uint32_t* __restrict__ out,
              int64_t n4, int F, int nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const float4 v = X[i];
  const int f = (int)((i * 4) % F);
  const uint32_t c0 = (uint8_t)bin_one(v.x, et + f, nb, F);
  const uint32_t c1 = (uint8_t)bin_one(v.y, et + f + 1, nb, F);
  const uint32_t c2 = (uint8_t)bin_one(v.z, et + f + 2, nb, F);
  const uint32_t c3 = (uint8_t)bin_one(v.w, et + f + 3, nb, F);
  out[i] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
}

// Tiled whole-matrix binning (F % 4 == 0, 16-byte aligned rows): a
// workgroup takes kBinFeat = 256 features x kBinRows rows with the tile's
// edges in LDS.  Lane l of every wave owns features 4l .. 4l+3 of the tile
// and reads them as one float4, so a wave reads 1 KB of one row per load and
// writes 256 contiguous code bytes; wave w takes rows w, w + 4, ...  Edge k
// of feature 4l + q is LDS word (q * kp + k) * 64 + l: its bank is fixed by
// the lane whatever edge it probes, so the 64 lanes hit each bank exactly
// twice.  A feature whose edges are non-decreasing and NaN-free takes a
// branch-free binary search for the number of edges <= x, its edges padded
// with NaN to kp = 2^m - 1 entries (x >= NaN is false, so no probe passes the
// last edge; NaN x gives 0); the probes of a thread's 16 searches (4 rows x 4
// features) are issued together, then compared.  This equals the
// reference's early-exit scan for such edges; any other feature takes that
// scan as a running AND.
constexpr int kBinRows = 256;
constexpr int kBinFeat = 256;
constexpr int kBinWGs = 8192;  // most workgroups of a launch (each loops over row blocks)

__host__ __device__ __forceinline__ int bin_pad(int nb) {
  int top = 1;
  while (2 * top <= nb) top *= 2;
  return 2 * top - 1;
}

// The edges, then one flag per feature.
__host__ __forceinline__ size_t bin_tile_lds(int nb) {
  return ((size_t)4 * bin_pad(nb) * 64 + kBinFeat) * sizeof(float);
}

// Edges of one feature: edge k at ef[k * 64].
__device__ __forceinline__ uint32_t code_scan(const float* __restrict__ ef, int nb, float x) {
  uint32_t a = 1, c = 0;
  for (int k = 0; k < nb; ++k) {
    a &= (uint32_t)(x >= ef[k * 64]);
    c += a;
  }
  return c;
}

__device__ __forceinline__ uint32_t code_search(const float* __restrict__ ef, int top, float x) {
  int lo = 0;
  for (int step = top; step > 0; step >>= 1) lo = (x >= ef[(lo + step - 1) * 64]) ? lo + step : lo;
  return (uint32_t)lo;
}

__global__ void __launch_bounds__(256)
k_bin_tile(const float4* __restrict__ X, const float* __restrict__ et, uint32_t* __restrict__ out, int64_t n,
           int F, int nb) {
  extern __shared__ float s_e[];  // no static LDS: s_e is at address 0
  // Workgroup b takes feature group b % n_groups of row blocks b / n_groups,
  // b / n_groups + gridDim.x / n_groups, ... (gridDim.x is a multiple of
  // n_groups): the groups of a row block run side by side, so each row's
  // bytes are read (and its codes written) by neighbouring workgroups at
  // about the same time, and each workgroup loads its edges once.
  const int n_groups = (F + kBinFeat - 1) / kBinFeat;
  const int f0 = (int)(blockIdx.x % n_groups) * kBinFeat;
  const int64_t n_rblk = (n + kBinRows - 1) / kBinRows;
  const int64_t rstride = gridDim.x / n_groups;
  const int nf = min(kBinFeat, F - f0);  // a multiple of 4
  const int kp = bin_pad(nb);            // >= nb; probes stay below it
  const int top = (kp + 1) >> 1;
  int* s_sorted = reinterpret_cast<int*>(s_e + 4 * kp * 64);
  for (int w = threadIdx.x; w < 4 * kp * 64; w += 256) {
    const int l = w & 63, qk = w >> 6;  // word w: lane l, entry qk = q * kp + k
    const int q = qk / kp, k = qk - q * kp;
    const int f = 4 * l + q;
    s_e[w] = (f < nf && k < nb) ? et[(int64_t)k * F + f0 + f] : __builtin_nanf("");
  }
  __syncthreads();
  {  // is feature t's edge list non-decreasing and NaN-free?
    const int l = threadIdx.x >> 2, q = threadIdx.x & 3;
    const float* ef = s_e + q * kp * 64 + l;
    bool ok = ef[0] == ef[0];
    for (int k = 1; k < nb; ++k) ok = ok && (ef[k * 64] >= ef[(k - 1) * 64]);  // false for NaN
    s_sorted[threadIdx.x] = ok;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (4 * lane >= nf) return;
  // LDS bytes: edge k of feature 4 lane + q at qb[q] + 256 k, so a probe is
  // one ds_read with the step in its offset.
  const char* sb = reinterpret_cast<const char*>(s_e);
  const int qb[4] = {lane * 4, lane * 4 + kp * 256, lane * 4 + 2 * kp * 256, lane * 4 + 3 * kp * 256};
  const bool all_sorted = s_sorted[4 * lane] && s_sorted[4 * lane + 1] && s_sorted[4 * lane + 2] &&
                          s_sorted[4 * lane + 3];
  const int64_t F4 = F >> 2;
  for (int64_t rblk = blockIdx.x / n_groups; rblk < n_rblk; rblk += rstride) {
  // Rows: wave wv takes rows row0 + 4 i, i = 0 .. kBinRows/4 - 1; a row's
  // address is a wave-uniform base plus the lane's float4 (or code word).
  const int64_t row0 = rblk * kBinRows + wv;
  const bool full = rblk * kBinRows + kBinRows <= n;  // uniform: every row of the block exists
  constexpr int RB = 4;  // rows per thread whose loads are in flight together
  constexpr int NI = kBinRows / 4 / RB;
  static_assert(kBinRows % (4 * RB) == 0, "row block");
  auto row_ok = [&](int i) { return full || row0 + 4 * i < n; };
  auto load_rows = [&](int i0, float4* v) {
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const float4* xb = X + (row0 + 4 * (i0 + j)) * F4 + (f0 >> 2);
      v[j] = row_ok(i0 + j) ? xb[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  // Software pipeline: the next RB rows' loads are issued before this RB's
  // searches, so they are in flight meanwhile.
  float4 vn[RB];
  load_rows(0, vn);
  for (int it = 0; it < NI; ++it) {
    float4 v[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) v[j] = vn[j];
    if (it + 1 < NI) load_rows((it + 1) * RB, vn);  // uniform
    uint32_t code[RB][4];
    if (all_sorted) {
      const char* lo[RB][4];  // the current lower bound's edge (256 bytes per edge passed)
#pragma unroll
      for (int j = 0; j < RB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) lo[j][q] = sb + qb[q];
#pragma unroll
      for (int sbit = 6; sbit >= 0; --sbit) {  // steps 64 .. 1, those up to top
        const int step = 1 << sbit;
        if (step > top) continue;  // uniform
        float e[RB][4];  // all probes of the step issued before any compare
#pragma unroll
        for (int j = 0; j < RB; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) e[j][q] = *reinterpret_cast<const float*>(lo[j][q] + (step - 1) * 256);
#pragma unroll
        for (int j = 0; j < RB; ++j) {
          const float xs[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
          for (int q = 0; q < 4; ++q) lo[j][q] += (xs[q] >= e[j][q]) ? step * 256 : 0;
        }
      }
#pragma unroll
      for (int j = 0; j < RB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) code[j][q] = (uint32_t)(lo[j][q] - (sb + qb[q])) >> 8;
    } else {
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const float xs[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float* ef = s_e + q * kp * 64 + lane;
          code[j][q] = s_sorted[4 * lane + q] ? code_search(ef, top, xs[q]) : code_scan(ef, nb, xs[q]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int i = it * RB + j;
      uint32_t* ob = out + (row0 + 4 * i) * F4 + (f0 >> 2);
      if (row_ok(i)) ob[lane] = code[j][0] | (code[j][1] << 8) | (code[j][2] << 16) | (code[j][3] << 24);
    }
  }
  }
}

__global__ void __launch_bounds__(256)
k_bin_matrix1(const float* __restrict__ X, const float* __restrict__ et, int8_t* __restrict__ out, int64_t n,
              int F, int nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = (int)(i % F);
  out[i] = bin_one(X[i], et + f, nb, F);
}

// Quantile edges of column-sorted data: sorted [n, ncol] (each column
// ascending, NaN last) -> edges [ncol, B-1] float32.  Same float32 steps as
// torch.quantile(col, torch.linspace(0, 1, B+1)[1:-1]) with linear
// interpolation (warpgbm/core.py:1271-1278): level q_j by one fused
// multiply-add as torch.linspace forms it, rank = q*(n-1), lo/hi = floor/
// ceil(rank), w = rank - lo, and the lerp with one rounding per branch as
// torch's lerp evaluates it: w < 0.5 ? fma(w, b-a, a) : fma(-(b-a), 1-w, b).
// Every other operation is rounded separately (built with -ffp-contract=off).
__global__ void __launch_bounds__(256)
k_edges_from_sorted(const float* __restrict__ sorted, int64_t n, int ncol, int num_bins,
                    float* __restrict__ edges, int edge_row_stride) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ne = num_bins - 1;
  if (t >= (int64_t)ncol * ne) return;
  const int c = (int)(t / ne);
  const int j = (int)(t - (int64_t)c * ne) + 1;  // interior level index 1..B-1
  const int steps = num_bins + 1;
  const float step = 1.0f / (float)(steps - 1);
  const int half = steps / 2;
  const float q = (j < half) ? fmaf(step, (float)j, 0.0f) : fmaf(-step, (float)(steps - 1 - j), 1.0f);
  float e;
  if (isnan(sorted[(n - 1) * ncol + c])) {
    e = __builtin_nanf("");
  } else {
    const float rank = q * (float)(n - 1);
    const float lo_f = floorf(rank), hi_f = ceilf(rank);
    // (float)(n-1) rounds up above 2^24 rows: keep both ranks in the column
    const int64_t lo = min((int64_t)lo_f, n - 1), hi = min((int64_t)hi_f, n - 1);
    const float w = rank - lo_f;
    const float a = sorted[lo * ncol + c];
    const float b = sorted[hi * ncol + c];
    const float d = b - a;
    e = (w < 0.5f) ? fmaf(w, d, a) : fmaf(-d, 1.0f - w, b);
  }
  edges[(int64_t)c * edge_row_stride + (j - 1)] = e;
}

}  // namespace wgbm

using namespace wgbm;

extern "C" int wgbm_edges_from_sorted(const float* sorted, int64_t n, int32_t ncol, int32_t num_bins, float* edges,
                                      int32_t edge_row_stride, hipStream_t stream) {
  WGBM_REQUIRE(n >= 1 && ncol >= 0 && num_bins >= 2 && num_bins <= 127 && edge_row_stride >= num_bins - 1,
               "bad quantile shape");
  const int64_t total = (int64_t)ncol * (num_bins - 1);
  if (total == 0) return kOk;
  hipLaunchKernelGGL(k_edges_from_sorted, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, sorted, n,
                     (int)ncol, (int)num_bins, edges, (int)edge_row_stride);
  WGBM_LAUNCHED("k_edges_from_sorted");
  return kOk;
}

extern "C" int wgbm_bin_column(const float* x, int64_t n, const float* bin_edges, int32_t n_edges, int8_t* out,
                               hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n_edges >= 0 && n_edges <= 127, "bad binning shape");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_bin_column, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, x, bin_edges, out, n,
                     (int)n_edges);
  WGBM_LAUNCHED("k_bin_column");
  return kOk;
}

extern "C" size_t wgbm_bin_matrix_workspace_bytes(int32_t n_feat, int32_t n_edges) {
  return (size_t)n_feat * (size_t)(n_edges > 0 ? n_edges : 1) * sizeof(float);
}

// X [n, F] row-major float32, edges [F, n_edges] float32 -> out [n, F] int8.
extern "C" int wgbm_bin_matrix(const float* X, int64_t n, int32_t F, const float* edges, int32_t n_edges,
                               int8_t* out, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && F >= 0 && n_edges >= 0 && n_edges <= 127, "bad binning shape");
  WGBM_REQUIRE(workspace_bytes >= wgbm_bin_matrix_workspace_bytes(F, n_edges), "binning workspace too small");
  const int64_t total = n * (int64_t)F;
  if (total == 0) return kOk;
  float* et = static_cast<float*>(workspace);
  if (n_edges > 0) {
    hipLaunchKernelGGL(k_transpose_edges, dim3((unsigned)ceil_div((int64_t)F * n_edges, 256)), dim3(256), 0,
                       stream, edges, et, (int)F, (int)n_edges);
    WGBM_LAUNCHED("k_transpose_edges");
  }
  const bool vec = (F % 4 == 0) && ((uintptr_t)X % 16 == 0) && ((uintptr_t)out % 4 == 0);
  if (vec && n_edges > 0) {
    const size_t lds = bin_tile_lds(n_edges);
    if (lds > 64 * 1024) {  // opt in once per device (up to 127 KB at 127 edges)
      static bool done[64] = {};
      int dev = 0;
      WGBM_HIP(hipGetDevice(&dev));
      if (dev < 0 || dev >= 64 || !done[dev]) {
        WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bin_tile),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bin_tile_lds(127)));
        if (dev >= 0 && dev < 64) done[dev] = true;
      }
    }
    const int64_t ng = ceil_div(F, kBinFeat);
    const int64_t per = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBinRows), kBinWGs / ng));
    const dim3 grid((unsigned)(per * ng));
    hipLaunchKernelGGL(k_bin_tile, grid, dim3(256), lds, stream, reinterpret_cast<const float4*>(X), et,
                       reinterpret_cast<uint32_t*>(out), n, (int)F, (int)n_edges);
    WGBM_LAUNCHED("k_bin_tile");
  } else if (vec) {
    const int64_t n4 = total / 4;
    hipLaunchKernelGGL(k_bin_matrix4, dim3((unsigned)ceil_div(n4, 256)), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(X), et, reinterpret_cast<uint32_t*>(out), n4, (int)F,
                       (int)n_edges);
    WGBM_LAUNCHED("k_bin_matrix4");
  } else {
    hipLaunchKernelGGL(k_bin_matrix1, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, X, et, out,
                       total, (int)F, (int)n_edges);
    WGBM_LAUNCHED("k_bin_matrix1");
  }
  return kOk;
}
