// Forest traversal (gfx950).
//
// 1. wgbm_predict_forest: the reference op on the reference's flattened
//    forest [T, max_nodes, 6] = (feature, bin, left, right, is_leaf, value)
//    (warpgbm/cuda/predict.cu:18-44, layout warpgbm/core.py:2283-2328).
//    One thread per row adds lr*leaf of every tree IN TREE ORDER (no
//    atomics), so results are deterministic and equal, bit for bit, to the
//    scores training accumulates.
// 2. A compiled "heap" forest for serving: every tree padded to a complete
//    binary tree of depth D (children of heap node i are 2i+1, 2i+2; a leaf
//    above depth D becomes pass-through nodes that always go left), internal
//    nodes packed as (feature << 8 | bin) in 32 bits and leaves stored
//    already multiplied by the learning rate.  Traversal is D branch-free
//    steps.  A workgroup stages the codes of 64 rows in LDS; wave q walks
//    trees q, q+4, ... of its tree block for those 64 rows; partial sums
//    are combined in a fixed order (tree block, then wave), so a row's
//    result does not depend on which other rows share the launch.
#include "common.h"
#include <mutex>

namespace wgbm {

__global__ void __launch_bounds__(256)
k_predict_ref(const int8_t* __restrict__ codes, int64_t N, int64_t F, const float* __restrict__ trees,
              int64_t T, int64_t max_nodes, float lr, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int8_t* row = codes + i * F;
  float acc = out[i];
  for (int64_t t = 0; t < T; ++t) {
    const float* tr = trees + t * max_nodes * 6;
    int64_t node = 0;
    for (int64_t step = 0; step <= max_nodes; ++step) {
      const float* nd = tr + node * 6;
      if (nd[4] > 0.5f) {
        acc = acc + lr * nd[5];
        break;
      }
      const int64_t feat = (int64_t)nd[0];
      const int bin = (int)nd[1];
      const int64_t nxt = (feat >= 0 && feat < F && (int)row[feat] <= bin) ? (int64_t)nd[2] : (int64_t)nd[3];
      if (!(nxt > 0 && nxt < max_nodes)) break;  // malformed tree: contributes nothing
      node = nxt;
    }
  }
  out[i] = acc;
}

// --------------------------------------------------------------- heap forest
struct HeapForest {
  const uint32_t* nodes;   // [T][2^D - 1]
  const float* leaves;     // [T][2^D], lr already applied
  int32_t T;
  int32_t depth;
  int32_t tree_block;      // trees per block (fixed per forest)
};

constexpr int kRowsPerWG = 64;

// Partial sums for (row tile, class, tree block): out_part[(cls*n_blocks +
// blk) * n + row] = sum over the block's trees.  Wave q of the workgroup
// adds trees q, q+4, q+8, ... of the block in that order, then the four wave
// sums are combined in wave order, so the result does not depend on how the
// traversal is scheduled.
//
// NODES_LDS: the block's nodes are staged in LDS kSubTrees trees at a time
// (every wave walks kSubTrees/4 trees of a sub-block at once, node reads
// from LDS); otherwise node reads go through the L1/L2 (the whole forest is
// a few MB) with INFLIGHT trees per lane at once.
constexpr int kSubTrees = 16;

template <bool NODES_LDS, int INFLIGHT>
__global__ void __launch_bounds__(256)
k_predict_heap(const int8_t* __restrict__ codes, int64_t n, int32_t F, HeapForest hf,
               const int32_t* __restrict__ class_first, int32_t n_classes, int32_t n_blocks,
               float* __restrict__ part) {
  extern __shared__ int8_t s_dyn[];  // [64][F] codes, then (NODES_LDS) [kSubTrees][internal] nodes
  __shared__ float s_sum[4][kRowsPerWG];
  const int64_t row0 = (int64_t)blockIdx.x * kRowsPerWG;
  const int cls = blockIdx.y / n_blocks;
  const int blk = blockIdx.y - cls * n_blocks;
  const int nrows = (int)min((int64_t)kRowsPerWG, n - row0);
  int8_t* s_codes = s_dyn;
  // stage codes (row-major, contiguous rows) with 4-byte loads where aligned
  const int64_t bytes = (int64_t)nrows * F;
  const int8_t* src = codes + row0 * F;
  if (((uintptr_t)src & 3) == 0) {
    const int64_t w4 = bytes >> 2;
    for (int64_t q = threadIdx.x; q < w4; q += 256)
      reinterpret_cast<uint32_t*>(s_codes)[q] = reinterpret_cast<const uint32_t*>(src)[q];
    for (int64_t q = (w4 << 2) + threadIdx.x; q < bytes; q += 256) s_codes[q] = src[q];
  } else {
    for (int64_t q = threadIdx.x; q < bytes; q += 256) s_codes[q] = src[q];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t_lo = class_first[cls] + blk * hf.tree_block;
  const int t_hi = min(t_lo + hf.tree_block, class_first[cls + 1]);
  const int internal = (1 << hf.depth) - 1;
  const int8_t* myrow = s_codes + (int64_t)(lane < nrows ? lane : 0) * F;
  float acc = 0.0f;
  if constexpr (NODES_LDS) {
    uint32_t* s_nodes = reinterpret_cast<uint32_t*>(s_dyn + (((int64_t)kRowsPerWG * F + 15) & ~(int64_t)15));
    for (int t0 = t_lo; t0 < t_hi; t0 += kSubTrees) {
      const int nt = min(kSubTrees, t_hi - t0);
      __syncthreads();  // previous sub-block's nodes are no longer read (and codes are staged)
      const uint32_t* g = hf.nodes + (int64_t)t0 * internal;
      for (int q = threadIdx.x; q < nt * internal; q += 256) s_nodes[q] = g[q];
      __syncthreads();
      constexpr int PER = kSubTrees / 4;  // trees of the sub-block per wave
      int at[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) at[k] = 0;
      for (int d = 0; d < hf.depth; ++d) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
          const int tt = wv + 4 * k;  // tree within the sub-block
          if (tt < nt) {
            const uint32_t x = s_nodes[tt * internal + at[k]];
            at[k] = 2 * at[k] + 1 + ((int)myrow[x >> 8] > (int)(x & 0xff));
          }
        }
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int tt = wv + 4 * k;
        if (tt < nt) acc = acc + hf.leaves[(int64_t)(t0 + tt) * (internal + 1) + (at[k] - internal)];
      }
    }
  } else {
    __syncthreads();
    int t = t_lo + wv;
    for (; t + 4 * (INFLIGHT - 1) < t_hi; t += 4 * INFLIGHT) {
      int at[INFLIGHT];
#pragma unroll
      for (int k = 0; k < INFLIGHT; ++k) at[k] = 0;
      for (int d = 0; d < hf.depth; ++d) {
#pragma unroll
        for (int k = 0; k < INFLIGHT; ++k) {
          const uint32_t x = hf.nodes[(int64_t)(t + 4 * k) * internal + at[k]];
          at[k] = 2 * at[k] + 1 + ((int)myrow[x >> 8] > (int)(x & 0xff));
        }
      }
#pragma unroll
      for (int k = 0; k < INFLIGHT; ++k)
        acc = acc + hf.leaves[(int64_t)(t + 4 * k) * (internal + 1) + (at[k] - internal)];
    }
    for (; t < t_hi; t += 4) {
      const uint32_t* n0 = hf.nodes + (int64_t)t * internal;
      int a = 0;
      for (int d = 0; d < hf.depth; ++d) {
        const uint32_t x = n0[a];
        a = 2 * a + 1 + ((int)myrow[x >> 8] > (int)(x & 0xff));
      }
      acc = acc + hf.leaves[(int64_t)t * (internal + 1) + (a - internal)];
    }
  }
  s_sum[wv][lane] = acc;
  __syncthreads();
  if (threadIdx.x < nrows) {
    const int r = threadIdx.x;
    const float v = ((s_sum[0][r] + s_sum[1][r]) + s_sum[2][r]) + s_sum[3][r];
    part[((int64_t)cls * n_blocks + blk) * n + row0 + r] = v;
  }
}

// Tree-order traversal: one lane per row adds base[cls] and then the
// class's trees one by one in tree order, exactly as training accumulates a
// row's score and as the reference op adds its trees, so predicting the
// training rows reproduces the training scores bit for bit.  INFLIGHT trees
// are walked at once (their leaves are still added in order).  A one-wave
// workgroup stages its 64 rows' codes in LDS.
template <int INFLIGHT>
__global__ void __launch_bounds__(64)
k_predict_heap_ordered(const int8_t* __restrict__ codes, int64_t n, int32_t F, HeapForest hf,
                       const int32_t* __restrict__ class_first, int32_t n_classes, const float* __restrict__ base,
                       float* __restrict__ out) {
  extern __shared__ int8_t s_codes[];  // [64][F]
  const int64_t row0 = (int64_t)blockIdx.x * kRowsPerWG;
  const int cls = blockIdx.y;
  const int nrows = (int)min((int64_t)kRowsPerWG, n - row0);
  const int64_t bytes = (int64_t)nrows * F;
  const int8_t* src = codes + row0 * F;
  if (((uintptr_t)src & 3) == 0) {
    const int64_t w4 = bytes >> 2;
    for (int64_t q = threadIdx.x; q < w4; q += 64)
      reinterpret_cast<uint32_t*>(s_codes)[q] = reinterpret_cast<const uint32_t*>(src)[q];
    for (int64_t q = (w4 << 2) + threadIdx.x; q < bytes; q += 64) s_codes[q] = src[q];
  } else {
    for (int64_t q = threadIdx.x; q < bytes; q += 64) s_codes[q] = src[q];
  }
  __syncthreads();
  const int lane = threadIdx.x;
  const int8_t* myrow = s_codes + (int64_t)(lane < nrows ? lane : 0) * F;
  const int internal = (1 << hf.depth) - 1;
  const int t1 = class_first[cls + 1];
  float acc = base[cls];
  int t = class_first[cls];
  for (; t + INFLIGHT <= t1; t += INFLIGHT) {
    int at[INFLIGHT];
#pragma unroll
    for (int k = 0; k < INFLIGHT; ++k) at[k] = 0;
    for (int d = 0; d < hf.depth; ++d) {
#pragma unroll
      for (int k = 0; k < INFLIGHT; ++k) {
        const uint32_t x = hf.nodes[(int64_t)(t + k) * internal + at[k]];
        at[k] = 2 * at[k] + 1 + ((int)myrow[x >> 8] > (int)(x & 0xff));
      }
    }
#pragma unroll
    for (int k = 0; k < INFLIGHT; ++k) acc = acc + hf.leaves[(int64_t)(t + k) * (internal + 1) + (at[k] - internal)];
  }
  for (; t < t1; ++t) {
    const uint32_t* n0 = hf.nodes + (int64_t)t * internal;
    int a = 0;
    for (int d = 0; d < hf.depth; ++d) {
      const uint32_t x = n0[a];
      a = 2 * a + 1 + ((int)myrow[x >> 8] > (int)(x & 0xff));
    }
    acc = acc + hf.leaves[(int64_t)t * (internal + 1) + (a - internal)];
  }
  if (lane < nrows) out[(row0 + lane) * n_classes + cls] = acc;
}

static int g_predict_mode = 0;  // 0 automatic, 1 global x2, 2 global x4, 3 nodes in LDS

// Allow up to 160 KB of dynamic LDS for a kernel, once per device.
template <bool NODES_LDS, int INFLIGHT>
static int predict_lds_opt_in(size_t bytes) {
  static bool done[64] = {};
  int dev = 0;
  WGBM_HIP(hipGetDevice(&dev));
  if (bytes <= 64 * 1024 || (dev >= 0 && dev < 64 && done[dev])) return kOk;
  WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_predict_heap<NODES_LDS, INFLIGHT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  if (dev >= 0 && dev < 64) done[dev] = true;
  return kOk;
}

// out[row*K + cls] = base[cls] + sum of block partials in block order.
__global__ void __launch_bounds__(256)
k_predict_heap_final(const float* __restrict__ part, int64_t n, int32_t n_classes, int32_t n_blocks,
                     const int32_t* __restrict__ blocks_of_class, const float* __restrict__ base,
                     float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n_classes) return;
  const int64_t r = i / n_classes;
  const int c = (int)(i - r * n_classes);
  float acc = base[c];
  const int nb = blocks_of_class[c];
  for (int b = 0; b < nb; ++b) acc = acc + part[((int64_t)c * n_blocks + b) * n + r];
  out[i] = acc;
}

// Build the heap forest from level-order node tables (heap ids): status
// 1 = split (feature, bin), 2 = leaf (value), 0 = absent.  Thread per (tree,
// heap slot of the padded tree).
__global__ void __launch_bounds__(256)
k_heap_compile(const int32_t* __restrict__ status, const int32_t* __restrict__ feat,
               const int32_t* __restrict__ bin, const float* __restrict__ value, int32_t src_nodes,
               int32_t T, int32_t depth, float lr, uint32_t* __restrict__ nodes, float* __restrict__ leaves) {
  const int internal = (1 << depth) - 1;
  const int total = 2 * internal + 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)T * total) return;
  const int t = (int)(i / total);
  const int h = (int)(i - (int64_t)t * total);
  const int32_t* st = status + (int64_t)t * src_nodes;
  // nearest ancestor-or-self that exists in the source table
  int a = h;
  while (a > 0 && (a >= src_nodes || st[a] == 0)) a = (a - 1) >> 1;
  if (h < internal) {
    uint32_t v = 127u;  // pass-through: feature 0, bin 127 -> always left
    if (a == h && st[a] == 1) v = ((uint32_t)feat[(int64_t)t * src_nodes + a] << 8) | (uint32_t)(bin[(int64_t)t * src_nodes + a] & 0xff);
    nodes[(int64_t)t * internal + h] = v;
  } else {
    float v = 0.0f;
    if (a < src_nodes && st[a] == 2) v = lr * value[(int64_t)t * src_nodes + a];
    leaves[(int64_t)t * (internal + 1) + (h - internal)] = v;
  }
}

extern "C" int wgbm_predict_forest(const int8_t* bin_indices, int64_t n_rows, int64_t n_feat, const float* trees,
                                   int64_t n_trees, int64_t max_nodes, float learning_rate, float* out,
                                   hipStream_t stream) {
  WGBM_REQUIRE(n_rows >= 0 && n_feat >= 0 && n_trees >= 0 && max_nodes >= 1, "bad forest shape");
  if (n_rows == 0 || n_trees == 0) return kOk;
  hipLaunchKernelGGL(k_predict_ref, dim3((unsigned)ceil_div(n_rows, 256)), dim3(256), 0, stream, bin_indices,
                     n_rows, n_feat, trees, n_trees, max_nodes, learning_rate, out);
  WGBM_LAUNCHED("k_predict_ref");
  return kOk;
}

extern "C" int wgbm_heap_compile(const int32_t* status, const int32_t* feat, const int32_t* bin, const float* value,
                                 int32_t src_nodes, int32_t n_trees, int32_t depth, float learning_rate,
                                 uint32_t* nodes, float* leaves, hipStream_t stream) {
  WGBM_REQUIRE(depth >= 1 && depth <= 16 && n_trees >= 0 && src_nodes >= 1, "bad heap forest shape");
  const int64_t total = (int64_t)n_trees * ((2 << depth) - 1);
  if (total == 0) return kOk;
  hipLaunchKernelGGL(k_heap_compile, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, status, feat, bin,
                     value, src_nodes, n_trees, depth, learning_rate, nodes, leaves);
  WGBM_LAUNCHED("k_heap_compile");
  return kOk;
}

// Traversal variant for tuning: 0 automatic, 1 nodes through the cache with
// two trees per lane in flight, 2 the same with four, 3 nodes staged in LDS.
extern "C" int wgbm_set_predict_mode(int32_t mode) {
  WGBM_REQUIRE(mode >= 0 && mode <= 3, "predict mode must be 0..3, got %d", mode);
  g_predict_mode = mode;
  return kOk;
}

extern "C" size_t wgbm_heap_predict_workspace_bytes(int64_t n_rows, int32_t n_classes, int32_t n_blocks) {
  return (size_t)n_rows * n_classes * n_blocks * sizeof(float);
}

// Traversal + fixed-order reduction of one batch into out, with the per-block
// partial sums in part ([n_classes * n_blocks][n_rows] floats).
static int heap_predict_launch(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                               const float* leaves, int32_t n_trees, int32_t depth, int32_t tree_block,
                               const int32_t* class_first, const int32_t* blocks_of_class, int32_t n_classes,
                               int32_t n_blocks, const float* base, float* out, float* part, hipStream_t stream) {
  HeapForest hf{nodes, leaves, n_trees, depth, tree_block};
  const dim3 grid((unsigned)ceil_div(n_rows, kRowsPerWG), (unsigned)(n_classes * n_blocks));
  const size_t code_lds = ((size_t)n_feat * kRowsPerWG + 15) & ~(size_t)15;
  const size_t node_lds = (size_t)kSubTrees * ((1u << depth) - 1) * 4;
  int mode = g_predict_mode;
  if (mode == 0) mode = 2;
  if (mode == 3 && code_lds + node_lds > 64 * 1024) mode = 2;  // sub-block does not fit beside the codes
  int rc = kOk;
  switch (mode) {
    case 1:
      rc = predict_lds_opt_in<false, 2>(code_lds);
      if (rc != kOk) return rc;
      hipLaunchKernelGGL((k_predict_heap<false, 2>), grid, dim3(256), code_lds, stream, codes, n_rows, n_feat, hf,
                         class_first, n_classes, n_blocks, part);
      break;
    case 3:
      hipLaunchKernelGGL((k_predict_heap<true, 1>), grid, dim3(256), code_lds + node_lds, stream, codes, n_rows,
                         n_feat, hf, class_first, n_classes, n_blocks, part);
      break;
    default:
      rc = predict_lds_opt_in<false, 4>(code_lds);
      if (rc != kOk) return rc;
      hipLaunchKernelGGL((k_predict_heap<false, 4>), grid, dim3(256), code_lds, stream, codes, n_rows, n_feat, hf,
                         class_first, n_classes, n_blocks, part);
  }
  WGBM_LAUNCHED("k_predict_heap");
  hipLaunchKernelGGL(k_predict_heap_final, dim3((unsigned)ceil_div(n_rows * n_classes, 256)), dim3(256), 0, stream,
                     part, n_rows, n_classes, n_blocks, blocks_of_class, base, out);
  WGBM_LAUNCHED("k_predict_heap_final");
  return kOk;
}

static int heap_predict_check(int64_t n_rows, int32_t n_feat, int32_t depth, int32_t tree_block, int32_t n_classes,
                              int32_t n_blocks) {
  WGBM_REQUIRE(n_rows >= 0 && n_feat >= 1 && depth >= 1 && depth <= 16 && tree_block >= 1 && n_classes >= 1 &&
                   n_blocks >= 1,
               "bad heap predict arguments");
  WGBM_REQUIRE((int64_t)n_feat * kRowsPerWG <= 96 * 1024, "too many features for the LDS row tile (%d)", n_feat);
  return kOk;
}

// Partial sums of wgbm_heap_predict_noalloc: one array per GPU, held for the
// life of the process.  Its host-side name is not a device address; the
// launcher uses the address hipGetSymbolAddress returns on each device.
constexpr int64_t kNoallocFloats = int64_t(4) << 20;  // 16 MiB
static __device__ __align__(256) float scratch[kNoallocFloats];
static float* g_scratch_ptr[64] = {};
static std::mutex g_scratch_mu;

// This device's address of `scratch`, looked up once per device (with that
// device current) on its first call.
static int noalloc_scratch(float** out) {
  int dev = 0;
  WGBM_HIP(hipGetDevice(&dev));
  WGBM_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    WGBM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    WGBM_REQUIRE(p != nullptr, "no device address for the predict scratch array");
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace wgbm

using namespace wgbm;

// Serve one batch of rows: out[row*K + k] = base[k] + sum over class-k trees.
// class_first [K+1]: tree ranges per class (trees of a class contiguous);
// blocks_of_class [K]: ceil(trees of class / tree_block).
extern "C" int wgbm_heap_predict(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                                 const float* leaves, int32_t n_trees, int32_t depth, int32_t tree_block,
                                 const int32_t* class_first, const int32_t* blocks_of_class, int32_t n_classes,
                                 int32_t n_blocks, const float* base, float* out, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream) {
  int rc = heap_predict_check(n_rows, n_feat, depth, tree_block, n_classes, n_blocks);
  if (rc != kOk) return rc;
  WGBM_REQUIRE(workspace_bytes >= wgbm_heap_predict_workspace_bytes(n_rows, n_classes, n_blocks),
               "predict workspace too small");
  if (n_rows == 0) return kOk;
  return heap_predict_launch(codes, n_rows, n_feat, nodes, leaves, n_trees, depth, tree_block, class_first,
                             blocks_of_class, n_classes, n_blocks, base, out, static_cast<float*>(workspace), stream);
}

// wgbm_heap_predict without a workspace: the partial sums live in this GPU's
// `scratch` array (no allocation of any kind).  Calls on one GPU share it,
// so calls that overlap in time on one GPU race.
extern "C" int wgbm_heap_predict_noalloc(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                                         const float* leaves, int32_t n_trees, int32_t depth, int32_t tree_block,
                                         const int32_t* class_first, const int32_t* blocks_of_class,
                                         int32_t n_classes, int32_t n_blocks, const float* base, float* out,
                                         hipStream_t stream) {
  int rc = heap_predict_check(n_rows, n_feat, depth, tree_block, n_classes, n_blocks);
  if (rc != kOk) return rc;
  if ((int64_t)n_classes * n_blocks > kNoallocFloats || n_rows > kNoallocFloats / ((int64_t)n_classes * n_blocks)) {
    set_error("predict without a workspace holds at most %lld partial sums (rows x classes x tree blocks); "
              "%lld x %d x %d exceeds it", (long long)kNoallocFloats, (long long)n_rows, n_classes, n_blocks);
    return kTooBig;
  }
  float* part = nullptr;
  rc = noalloc_scratch(&part);  // every call, including one with no rows, sets the device's address up
  if (rc != kOk) return rc;
  if (n_rows == 0) return kOk;
  return heap_predict_launch(codes, n_rows, n_feat, nodes, leaves, n_trees, depth, tree_block, class_first,
                             blocks_of_class, n_classes, n_blocks, base, out, part, stream);
}

// Tree-order predict of a heap forest (no workspace): out[row*K + k] =
// base[k] + the class-k trees' leaves added one by one in tree order.
extern "C" int wgbm_heap_predict_tree_order(const int8_t* codes, int64_t n_rows, int32_t n_feat,
                                            const uint32_t* nodes, const float* leaves, int32_t n_trees,
                                            int32_t depth, const int32_t* class_first, int32_t n_classes,
                                            const float* base, float* out, hipStream_t stream) {
  WGBM_REQUIRE(n_rows >= 0 && n_feat >= 1 && depth >= 1 && depth <= 16 && n_classes >= 1 && n_trees >= 0,
               "bad heap predict arguments");
  WGBM_REQUIRE((int64_t)n_feat * kRowsPerWG <= 96 * 1024, "too many features for the LDS row tile (%d)", n_feat);
  if (n_rows == 0) return kOk;
  HeapForest hf{nodes, leaves, n_trees, depth, 1};
  const size_t lds = (size_t)n_feat * kRowsPerWG;
  if (lds > 64 * 1024) {
    static bool done[64] = {};
    int dev = 0;
    WGBM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || !done[dev]) {
      WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_predict_heap_ordered<4>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
      if (dev >= 0 && dev < 64) done[dev] = true;
    }
  }
  hipLaunchKernelGGL((k_predict_heap_ordered<4>), dim3((unsigned)ceil_div(n_rows, kRowsPerWG), (unsigned)n_classes),
                     dim3(64), lds, stream, codes, n_rows, n_feat, hf, class_first, n_classes, base, out);
  WGBM_LAUNCHED("k_predict_heap_ordered");
  return kOk;
}
