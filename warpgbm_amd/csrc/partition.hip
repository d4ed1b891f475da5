// Row partitioning, leaf values, score updates, histogram subtraction and
// gradient/hessian kernels (gfx950).
//
// Positions of a tree level are grouped by node, and within a node by era
// (segment (node, era)).  A split moves every segment's rows into the left
// child (code <= bin) or the right child, each side keeping the parent's
// order (stable; warpgbm/core.py:1425-1427 compacts with a boolean mask),
// so children stay era-grouped.  Children occupy the parent's range: left
// child first, then right.
#include "common.h"
#include "wgbm_internal.h"
#include "wgbm_level.h"
#include "wgbm_stats.h"
#include "wgbm_plan.h"
#include "wgbm_cand.h"

namespace wgbm {

// ----------------------------------------------------------------- partition
__device__ __forceinline__ bool is_split(const PartArgs& a, int j) {
  return a.node_active[j] && a.dec_feat[j] >= 0;
}

__device__ __forceinline__ int split_col(const PartArgs& a, int j) {
  const int fl = a.dec_feat[j];
  return a.feat_idx ? a.feat_idx[fl] : fl;
}

__device__ __forceinline__ bool goes_left(const PartArgs& a, int row, int col, int bin) {
  if ((uint64_t)row >= (uint64_t)a.n_rows) return true;
  // Column-major copy when present: a split column is then n_rows contiguous
  // bytes (L2-resident) instead of one 128-byte line per row.
  const int8_t code = a.codes_t ? a.codes_t[(int64_t)col * a.n_rows + row] : a.codes[(int64_t)row * a.row_stride + col];
  return (int)code <= bin;
}

// Block bodies take the chunk as an argument: the kernels below run them one
// chunk per workgroup, and the batched engine's fused level launches run
// several of them side by side over disjoint workgroup ranges.
// Segment of a chunk: the plan's map when present, else a search.
__device__ __forceinline__ int chunk_segment(const PartArgs& a, int nseg, int chunk) {
  return a.chunk_seg ? a.chunk_seg[chunk] : upper_index(a.seg_first, nseg, chunk);
}

// Go-left flags of positions [c0, c1) of one split segment (column col,
// threshold bin) and their count in lbase[chunk].
__device__ __forceinline__ void part_count_rows(const PartArgs& a, int chunk, int c0, int c1, int col, int bin) {
  // U positions per thread at once: their row ids, then their codes, are
  // loaded together (two dependent round trips per U positions).
  constexpr int U = 4;
  int cnt = 0;
  for (int p0 = c0 + threadIdx.x; p0 < c1; p0 += U * blockDim.x) {
    int rows[U];
#pragma unroll
    for (int u = 0; u < U; ++u) rows[u] = a.perm_in[min(p0 + u * (int)blockDim.x, c1 - 1)];
    bool left[U];
#pragma unroll
    for (int u = 0; u < U; ++u) left[u] = goes_left(a, rows[u], col, bin);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = p0 + u * (int)blockDim.x;
      if (p < c1) {
        a.flag[p] = (uint8_t)left[u];  // the scatter reads this instead of the code
        cnt += left[u];
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o);
  __shared__ int s_c[4];
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) a.lbase[chunk] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

__device__ __forceinline__ void part_count_block(const PartArgs& a, int chunk) {
  const int nseg = *a.nseg_dev;
  if (chunk >= a.seg_first[nseg]) return;
  const int s = chunk_segment(a, nseg, chunk);
  const int k = chunk - a.seg_first[s];
  const int j = s / a.n_eras;
  const int c0 = a.seg_start[s] + k * a.chunk_rows;
  const int c1 = min(c0 + a.chunk_rows, a.seg_start[s] + max(a.seg_count[s], 0));
  part_count_rows(a, chunk, c0, c1, split_col(a, j), a.dec_bin[j]);
}

__global__ void __launch_bounds__(256)
k_part_count(PartArgs a) { part_count_block(a, blockIdx.x); }

// One workgroup: per-segment left totals and chunk offsets, child segment
// tables, smaller-child build list and sibling map.

constexpr int kPsStage = 16384;  // values per staged scan pass (two arrays of them in dynamic LDS)
constexpr int kPsLoads = 8;     // items a thread loads together in the gather loops
// Stage index of item j: one pad word after every 32, so that the threads of
// a wave, each scanning its own run of consecutive items, hit distinct banks.
__device__ __forceinline__ int ps_pad(int j) { return j + (j >> 5); }
constexpr int kPsStagePad = kPsStage + kPsStage / 32;

// Exclusive scan of two values per item over n items (every thread calls
// it): passes of kPsStage items are loaded into LDS (kPsLoads loads in flight
// per thread, coalesced), each thread scans kPsStage / blockDim consecutive
// ones, then put(i, prefix0, prefix1) stores coalesced.  Returns the totals.
template <class Get, class Put>
__device__ __forceinline__ int2 scan2_staged(int n, Get get, Put put, int* s0, int* s1) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int cpt = kPsStage / nt;
  int2 carry = make_int2(0, 0);
  for (int base = 0; base < n; base += kPsStage) {
    const int span = min(kPsStage, n - base);
    for (int k0 = 0; k0 < span; k0 += kPsLoads * nt) {
      int2 v[kPsLoads];
#pragma unroll
      for (int u = 0; u < kPsLoads; ++u) {
        const int k = k0 + u * nt + t;
        v[u] = k < span ? get(base + k) : make_int2(0, 0);
      }
#pragma unroll
      for (int u = 0; u < kPsLoads; ++u) {
        const int k = k0 + u * nt + t;
        if (k < span) { s0[ps_pad(k)] = v[u].x; s1[ps_pad(k)] = v[u].y; }
      }
    }
    __syncthreads();
    int a0 = 0, a1 = 0;
    for (int i = 0; i < cpt; ++i) {
      const int j = t * cpt + i;
      if (j < span) { a0 += s0[ps_pad(j)]; a1 += s1[ps_pad(j)]; }
    }
    int tot0, tot1;
    int r0 = carry.x + block_exscan(a0, &tot0);
    int r1 = carry.y + block_exscan(a1, &tot1);
    for (int i = 0; i < cpt; ++i) {
      const int j = t * cpt + i;
      if (j < span) {
        const int x0 = s0[ps_pad(j)], x1 = s1[ps_pad(j)];
        s0[ps_pad(j)] = r0; s1[ps_pad(j)] = r1;
        r0 += x0; r1 += x1;
      }
    }
    __syncthreads();
    for (int k = t; k < span; k += nt) put(base + k, s0[ps_pad(k)], s1[ps_pad(k)]);
    carry.x += tot0;
    carry.y += tot1;
    __syncthreads();  // the next pass reuses the stage
  }
  return carry;
}

// One workgroup: per-chunk destinations, child tables and the build list.
// Chunks of a segment are contiguous, so one scan over all chunks gives
// every chunk its left rows before it within its segment.  Scans go through
// LDS in passes; the gather loops issue kPsLoads items' loads together.
__global__ void __launch_bounds__(1024)
k_part_scan(PartArgs a) {
  extern __shared__ int s_stage[];  // [2][kPsStagePad]
  int* s_v0 = s_stage;
  int* s_v1 = s_stage + kPsStagePad;
  const int E = a.n_eras;
  const int nseg = a.n_nodes * E;
  const int t = threadIdx.x, nt = blockDim.x;
  const int total = a.seg_first[nseg];
  // A: exclusive scan of the per-chunk left counts.
  const int2 A = scan2_staged(
      total, [&](int c) { return make_int2(a.lbase[c], 0); }, [&](int c, int x, int) { a.cscan[c] = x; }, s_v0, s_v1);
  if (t == 0) a.cscan[total] = A.x;
  __syncthreads();
  // B: child tables.  Exclusive scans over all segments of the left and
  // right counts give each era's offset inside its node's children
  // (segP[0..nseg] left, segP[nseg+1..2nseg+1] right).
  int* PL = a.segP;
  int* PR = a.segP + nseg + 1;
  const int2 B = scan2_staged(
      nseg,
      [&](int s) {
        if (!is_split(a, s / E)) return make_int2(0, 0);
        const int L = a.cscan[a.seg_first[s + 1]] - a.cscan[a.seg_first[s]];
        return make_int2(L, max(a.seg_count[s], 0) - L);
      },
      [&](int s, int pl, int pr) { PL[s] = pl; PR[s] = pr; }, s_v0, s_v1);
  if (t == 0) { PL[nseg] = B.x; PR[nseg] = B.y; }
  __syncthreads();
  for (int j = t; j < a.n_nodes; j += nt) {
    const int l = 2 * j, r = 2 * j + 1;
    if (!is_split(a, j)) {
      a.c_active[l] = 0; a.c_active[r] = 0;
      a.c_node_start[l] = 0; a.c_node_start[r] = 0;
      a.c_node_count[l] = 0; a.c_node_count[r] = 0;
      continue;
    }
    const int start = a.node_start[j];
    const int Lt = PL[(j + 1) * E] - PL[j * E];
    a.c_active[l] = 1; a.c_active[r] = 1;
    a.c_node_start[l] = start; a.c_node_count[l] = Lt;
    a.c_node_start[r] = start + Lt; a.c_node_count[r] = a.node_count[j] - Lt;
  }
  for (int s0 = t; s0 < nseg; s0 += kPsLoads * nt) {
    int pl0[kPsLoads], pl1[kPsLoads], pr0[kPsLoads], pr1[kPsLoads], nl0[kPsLoads], nl1[kPsLoads], nr0[kPsLoads],
        st[kPsLoads];
    bool sp[kPsLoads];
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int s = min(s0 + u * nt, nseg - 1);
      const int j = s / E;
      sp[u] = is_split(a, j);
      pl0[u] = PL[s]; pl1[u] = PL[s + 1]; pr0[u] = PR[s]; pr1[u] = PR[s + 1];
      nl0[u] = PL[j * E]; nl1[u] = PL[(j + 1) * E]; nr0[u] = PR[j * E];
      st[u] = a.node_start[j];
    }
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int s = s0 + u * nt;
      if (s >= nseg) break;
      const int j = s / E, e = s - j * E;
      const int l = 2 * j, r = 2 * j + 1;
      if (!sp[u]) {
        a.c_seg_start[l * E + e] = 0; a.c_seg_count[l * E + e] = 0;
        a.c_seg_start[r * E + e] = 0; a.c_seg_count[r * E + e] = 0;
        continue;
      }
      const int Lt = nl1[u] - nl0[u];
      a.c_seg_start[l * E + e] = st[u] + pl0[u] - nl0[u];
      a.c_seg_count[l * E + e] = pl1[u] - pl0[u];
      a.c_seg_start[r * E + e] = st[u] + Lt + pr0[u] - nr0[u];
      a.c_seg_count[r * E + e] = pr1[u] - pr0[u];
    }
  }
  __syncthreads();
  // C: absolute destination bases per chunk (only split segments own chunks).
  for (int c0 = t; c0 < total; c0 += kPsLoads * nt) {
    int sg[kPsLoads], fi[kPsLoads], cs[kPsLoads], cf[kPsLoads], ls[kPsLoads], rs[kPsLoads];
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int c = min(c0 + u * nt, total - 1);
      sg[u] = chunk_segment(a, nseg, c);
      cs[u] = a.cscan[c];
    }
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int j = sg[u] / E, e = sg[u] - j * E;
      fi[u] = a.seg_first[sg[u]];
      ls[u] = a.c_seg_start[(2 * j) * E + e];
      rs[u] = a.c_seg_start[(2 * j + 1) * E + e];
    }
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) cf[u] = a.cscan[fi[u]];
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int c = c0 + u * nt;
      if (c >= total) break;
      const int lb = cs[u] - cf[u];
      a.lbase[c] = lb + ls[u];
      a.rbase[c] = (c - fi[u]) * a.chunk_rows - lb + rs[u];
    }
  }
  // D: build list (smaller child, left on ties: warpgbm/core.py:1438) in node order.
  int rank0 = 0;
  for (int base = 0; base < a.n_nodes; base += nt) {
    const int j = base + t;
    const bool split = j < a.n_nodes && is_split(a, j);
    int tot;
    const int rank = rank0 + block_exscan(split ? 1 : 0, &tot);
    if (split) {
      const int l = 2 * j, r = 2 * j + 1;
      const int built = (a.c_node_count[l] <= a.c_node_count[r]) ? l : r;
      a.sib[3 * rank + 0] = j;
      a.sib[3 * rank + 1] = built;
      a.sib[3 * rank + 2] = built ^ 1;
    }
    rank0 += tot;
  }
  if (t == 0) {
    *a.n_split = rank0;
    *a.n_build = rank0 * E;
  }
  __syncthreads();
  const int ns = rank0;
  for (int q0 = t; q0 < ns * E; q0 += kPsLoads * nt) {
    int ch[kPsLoads], ss[kPsLoads], sc[kPsLoads];
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) ch[u] = a.sib[3 * (min(q0 + u * nt, ns * E - 1) / E) + 1];
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int q = min(q0 + u * nt, ns * E - 1);
      const int e = q - (q / E) * E;
      ss[u] = a.c_seg_start[ch[u] * E + e];
      sc[u] = a.c_seg_count[ch[u] * E + e];
    }
#pragma unroll
    for (int u = 0; u < kPsLoads; ++u) {
      const int q = q0 + u * nt;
      if (q >= ns * E) break;
      const int e = q - (q / E) * E;
      a.b_seg_start[q] = ss[u];
      a.b_seg_count[q] = sc[u];
      a.b_seg_slot[q] = ch[u] * E + e;
    }
  }
}

// k_part_scan with every table of the level in LDS: one round of loads, the
// scans in LDS, the outputs stored once.  Same outputs as k_part_scan (lbase
// and rbase, child tables, build list, sibling map); the launch picks it
// when the level has at most kScanLdsChunks chunks and kScanLdsSeg segments.
constexpr int kScanLdsChunks = 4096;
constexpr int kScanLdsSeg = 512;

__global__ void __launch_bounds__(1024)
k_part_scan_lds(PartArgs a) {
  __shared__ int s_cs[kScanLdsChunks + 1];   // exclusive scan of the chunks' left counts
  __shared__ int s_sf[kScanLdsSeg + 1];      // first chunk of each segment
  __shared__ int s_PL[kScanLdsSeg + 1], s_PR[kScanLdsSeg + 1];
  __shared__ int s_split[kScanLdsSeg];       // per node
  __shared__ int s_cst[2 * kScanLdsSeg];     // children's segment starts [child * E + e]
  __shared__ int s_ccount[2 * kScanLdsSeg];  // children's segment counts
  __shared__ int s_nc[2 * kScanLdsSeg];      // children's node counts
  __shared__ int s_built[kScanLdsSeg];       // built child of each split, in node order
  const int E = a.n_eras;
  const int nseg = a.n_nodes * E;
  const int t = threadIdx.x;
  const int total = a.seg_first[nseg];
  // Loads: segment and node tables, four chunk counts per thread.
  for (int s = t; s <= nseg; s += blockDim.x) s_sf[s] = a.seg_first[s];
  for (int j = t; j < a.n_nodes; j += blockDim.x) s_split[j] = is_split(a, j) ? 1 : 0;
  constexpr int CPT = kScanLdsChunks / 1024;  // chunks per thread
  int v[CPT];
  int mine = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = t * CPT + q;
    v[q] = c < total ? a.lbase[c] : 0;
    mine += v[q];
  }
  int tot;
  int run = block_exscan(mine, &tot);  // syncs: s_sf and s_split are complete
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = t * CPT + q;
    if (c <= total) s_cs[c] = run;
    run += v[q];
  }
  if (t == 0) s_cs[total] = tot;
  __syncthreads();
  // Left / right row counts per segment, scanned over all segments.
  int L = 0, R = 0;
  if (t < nseg && s_split[t / E]) {
    L = s_cs[s_sf[t + 1]] - s_cs[s_sf[t]];
    R = max(a.seg_count[t], 0) - L;
  }
  int tl, tr;
  const int el = block_exscan(L, &tl);
  const int er = block_exscan(R, &tr);
  if (t < nseg) { s_PL[t] = el; s_PR[t] = er; }
  if (t == 0) { s_PL[nseg] = tl; s_PR[nseg] = tr; }
  __syncthreads();
  // Child tables.
  for (int j = t; j < a.n_nodes; j += blockDim.x) {
    const int l = 2 * j, r = 2 * j + 1;
    if (!s_split[j]) {
      a.c_active[l] = 0; a.c_active[r] = 0;
      a.c_node_start[l] = 0; a.c_node_start[r] = 0;
      a.c_node_count[l] = 0; a.c_node_count[r] = 0;
      s_nc[l] = 0; s_nc[r] = 0;
      continue;
    }
    const int start = a.node_start[j];
    const int Lt = s_PL[(j + 1) * E] - s_PL[j * E];
    const int Rt = a.node_count[j] - Lt;
    a.c_active[l] = 1; a.c_active[r] = 1;
    a.c_node_start[l] = start; a.c_node_count[l] = Lt;
    a.c_node_start[r] = start + Lt; a.c_node_count[r] = Rt;
    s_nc[l] = Lt; s_nc[r] = Rt;
  }
  for (int s = t; s < nseg; s += blockDim.x) {
    const int j = s / E, e = s - j * E;
    const int l = 2 * j, r = 2 * j + 1;
    int ls = 0, lc = 0, rs = 0, rc = 0;
    if (s_split[j]) {
      const int start = a.node_start[j];
      const int Lt = s_PL[(j + 1) * E] - s_PL[j * E];
      ls = start + s_PL[s] - s_PL[j * E];
      lc = s_PL[s + 1] - s_PL[s];
      rs = start + Lt + s_PR[s] - s_PR[j * E];
      rc = s_PR[s + 1] - s_PR[s];
    }
    a.c_seg_start[l * E + e] = ls; a.c_seg_count[l * E + e] = lc;
    a.c_seg_start[r * E + e] = rs; a.c_seg_count[r * E + e] = rc;
    s_cst[l * E + e] = ls; s_ccount[l * E + e] = lc;
    s_cst[r * E + e] = rs; s_ccount[r * E + e] = rc;
  }
  __syncthreads();
  // Absolute destination bases per chunk (only split segments own chunks).
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = t * CPT + q;
    if (c >= total) break;
    const int s = upper_index(s_sf, nseg, c);
    const int j = s / E, e = s - j * E;
    const int first = s_sf[s];
    const int lb = s_cs[c] - s_cs[first];
    a.lbase[c] = lb + s_cst[(2 * j) * E + e];
    a.rbase[c] = (c - first) * a.chunk_rows - lb + s_cst[(2 * j + 1) * E + e];
  }
  // Build list (smaller child, left on ties: warpgbm/core.py:1438) in node order.
  const int sp = t < a.n_nodes ? s_split[t] : 0;
  int ns;
  const int rank = block_exscan(sp, &ns);
  if (sp) {
    const int l = 2 * t, r = 2 * t + 1;
    const int built = (s_nc[l] <= s_nc[r]) ? l : r;
    a.sib[3 * rank + 0] = t;
    a.sib[3 * rank + 1] = built;
    a.sib[3 * rank + 2] = built ^ 1;
    s_built[rank] = built;
  }
  if (t == 0) {
    *a.n_split = ns;
    *a.n_build = ns * E;
  }
  __syncthreads();
  for (int u = t; u < ns * E; u += blockDim.x) {
    const int q = u / E, e = u - q * E;
    const int c = s_built[q];
    a.b_seg_start[u] = s_cst[c * E + e];
    a.b_seg_count[u] = s_ccount[c * E + e];
    a.b_seg_slot[u] = c * E + e;
  }
}

static int part_scan_launch(const PartArgs& a, int64_t max_chunks, hipStream_t st) {
  const int64_t nseg = (int64_t)a.n_nodes * a.n_eras;
  if (max_chunks + 1 <= kScanLdsChunks && nseg <= kScanLdsSeg && a.n_nodes <= 1024) {
    hipLaunchKernelGGL(k_part_scan_lds, dim3(1), dim3(1024), 0, st, a);
    WGBM_LAUNCHED("k_part_scan_lds");
  } else {
    // One workgroup; its two stage arrays (132 KB) take most of a CU's LDS,
    // which needs the launch's dynamic LDS limit raised once per device.
    constexpr size_t lds = 2 * kPsStagePad * sizeof(int);
    static bool opted[64] = {};
    int dev = 0;
    WGBM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || !opted[dev]) {
      WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_part_scan),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      if (dev >= 0 && dev < 64) opted[dev] = true;
    }
    hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), lds, st, a);
    WGBM_LAUNCHED("k_part_scan");
  }
  return kOk;
}

__device__ __forceinline__ void part_scatter_block(const PartArgs& a, int chunk) {
  const int nseg = *a.nseg_dev;
  if (chunk >= a.seg_first[nseg]) return;
  const int s = chunk_segment(a, nseg, chunk);
  const int k = chunk - a.seg_first[s];
  const int c0 = a.seg_start[s] + k * a.chunk_rows;
  const int c1 = min(c0 + a.chunk_rows, a.seg_start[s] + max(a.seg_count[s], 0));
  __shared__ int s_l[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int lrun = a.lbase[chunk], rrun = a.rbase[chunk];
  StatAcc sl, sr;  // statistics of the chunk's left and right rows
  // Passes of 256 positions, U at a time: the U passes' loads are issued
  // together, then each pass ranks its rows and stores them.
  constexpr int U = 4;
  for (int base0 = c0; base0 < c1; base0 += U * 256) {
    int rows[U];
    float2 vs[U];
    bool lefts[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = min(base0 + u * 256 + (int)threadIdx.x, c1 - 1);
      rows[u] = a.perm_in[p];
      vs[u] = a.gh_in ? a.gh_in[p] : make_float2(0.0f, 0.0f);
      lefts[u] = a.flag[p] != 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int base = base0 + u * 256;
      if (base >= c1) break;  // uniform
      const int p = base + threadIdx.x;
      const bool ok = p < c1;
      const int row = rows[u];
      const float2 v = ok ? vs[u] : make_float2(0.0f, 0.0f);
      const bool left = ok && lefts[u];
      const uint64_t lm = __ballot(ok && left);
      const uint64_t rm = __ballot(ok && !left);
      if (lane == 0) s_l[wv] = __popcll(lm);
      __syncthreads();
      int lpre = 0, ltot = 0;
      for (int w = 0; w < 4; ++w) { const int c = s_l[w]; if (w < wv) lpre += c; ltot += c; }
      const int n_here = min(256, c1 - base);
      const int wave_lo = min(64 * wv, n_here);           // positions before this wave
      const int rpre = wave_lo - lpre;
      if (ok) {
        const int dst = left ? (lrun + lpre + lanes_below(lm)) : (rrun + rpre + lanes_below(rm));
        a.perm_out[dst] = row;
        if (a.gh_out) a.gh_out[dst] = v;
        if (left) sl.add(v.x, v.y);
        else sr.add(v.x, v.y);
      }
      lrun += ltot;
      rrun += n_here - ltot;
      __syncthreads();
    }
  }
  if (a.c_stats) {  // one partial per chunk and side, folded per child by k_child_stats
    sl.block_reduce();
    sr.block_reduce();
    if (threadIdx.x == 0) {
      sl.store(a.stat_part + ((int64_t)chunk * 2 + 0) * kStatWords);
      sr.store(a.stat_part + ((int64_t)chunk * 2 + 1) * kStatWords);
    }
  }
}

__global__ void __launch_bounds__(256)
k_part_scatter(PartArgs a) { part_scatter_block(a, blockIdx.x); }

// Child c of node j = c / 2 (side c % 2) folds the partials of j's chunks,
// which are contiguous: [seg_first[j*E], seg_first[(j+1)*E]).
__global__ void __launch_bounds__(256)
k_child_stats(PartArgs a) {
  const int c = blockIdx.x;
  const int j = c >> 1, side = c & 1;
  const int first = a.seg_first[j * a.n_eras], last = a.seg_first[(j + 1) * a.n_eras];
  StatAcc acc;
  for (int k = first + threadIdx.x; k < last; k += blockDim.x)
    acc.load_merge(a.stat_part + ((int64_t)k * 2 + side) * kStatWords);
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(a.c_stats + (int64_t)c * kStatWords);
}

PlanJob part_plan_job(const PartArgs& a) {
  PlanJob j{};
  j.src.mode = kPlanSplits;
  j.src.nseg = a.n_nodes * a.n_eras;
  j.src.nseg_out = a.nseg_dev;
  j.src.seg_count = a.seg_count;
  j.src.node_active = a.node_active;
  j.src.dec_feat = a.dec_feat;
  j.src.n_eras = a.n_eras;
  j.src.chunk_seg = a.chunk_seg;
  j.chunk_rows = a.chunk_rows;
  j.seg_first = a.seg_first;
  j.multi = a.multi;
  j.n_multi = a.n_multi;
  return j;
}

// Child statistics (workgroups 0 .. 2n-1, as k_child_stats) and the next
// level's histogram plan (last workgroup) in one launch.
__global__ void __launch_bounds__(1024)
k_child_stats_plan(PartArgs a, const PlanJob next) {
  const int c = blockIdx.x;
  if (c == 2 * a.n_nodes) {
    plan_body(next.src, next.chunk_rows, next.seg_first, next.seg_pfirst, next.multi, next.n_multi);
    return;
  }
  const int j = c >> 1, side = c & 1;
  const int first = a.seg_first[j * a.n_eras], last = a.seg_first[(j + 1) * a.n_eras];
  StatAcc acc;
  for (int k = first + threadIdx.x; k < last; k += blockDim.x)
    acc.load_merge(a.stat_part + ((int64_t)k * 2 + side) * kStatWords);
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(a.c_stats + (int64_t)c * kStatWords);
}

int partition_launch(const PartArgs& a, int64_t rows_max, hipStream_t st, bool planned, const PlanJob* next_hist) {
  const int nseg = a.n_nodes * a.n_eras;
  if (nseg == 0) return kOk;
  WGBM_REQUIRE(a.codes && a.node_active && a.dec_feat && a.dec_bin && a.seg_start && a.seg_count && a.node_start &&
                   a.node_count && a.perm_in && a.perm_out && a.flag && a.seg_first && a.nseg_dev && a.multi &&
                   a.n_multi && a.lbase && a.rbase && a.cscan && a.segL && a.segP && a.c_seg_start &&
                   a.c_seg_count && a.c_active && a.c_node_start && a.c_node_count && a.b_seg_start &&
                   a.b_seg_count && a.b_seg_slot && a.n_build && a.sib && a.n_split && a.chunk_rows > 0,
               "partition: missing table or scratch pointer");
  WGBM_REQUIRE(!a.c_stats || a.stat_part, "partition: child statistics need their scratch");
  if (!planned) {
    const PlanJob j = part_plan_job(a);
    plan_launch(j.src, j.chunk_rows, j.seg_first, j.multi, j.n_multi, st);
    WGBM_LAUNCHED("k_hist_plan(partition)");
  }
  const int64_t max_chunks = ceil_div(rows_max, a.chunk_rows) + nseg;
  hipLaunchKernelGGL(k_part_count, dim3((unsigned)max_chunks), dim3(256), 0, st, a);
  WGBM_LAUNCHED("k_part_count");
  const int rc = part_scan_launch(a, max_chunks, st);
  if (rc != kOk) return rc;
  hipLaunchKernelGGL(k_part_scatter, dim3((unsigned)max_chunks), dim3(256), 0, st, a);
  WGBM_LAUNCHED("k_part_scatter");
  if (a.c_stats && next_hist) {
    hipLaunchKernelGGL(k_child_stats_plan, dim3((unsigned)(2 * a.n_nodes + 1)), dim3(kPlanThreads), 0, st, a, *next_hist);
    WGBM_LAUNCHED("k_child_stats_plan");
  } else if (a.c_stats) {
    hipLaunchKernelGGL(k_child_stats, dim3((unsigned)(2 * a.n_nodes)), dim3(256), 0, st, a);
    WGBM_LAUNCHED("k_child_stats");
  }
  return kOk;
}

// Sum of a 128-bit integer over the 64 lanes of a wave (every lane gets it).
__device__ __forceinline__ I128 wave_sum128(I128 x) {
  for (int o = 32; o >= 1; o >>= 1) {
    I128 y;
    y.lo = ((unsigned long long)(uint32_t)__shfl_xor((int)(uint32_t)(x.lo >> 32), o) << 32) |
           (uint32_t)__shfl_xor((int)(uint32_t)x.lo, o);
    y.hi = (long long)(((unsigned long long)(uint32_t)__shfl_xor((int)(uint32_t)((unsigned long long)x.hi >> 32), o)
                        << 32) |
                       (uint32_t)__shfl_xor((int)(uint32_t)(unsigned long long)x.hi, o));
    x.add(y);
  }
  return x;
}

// ----------------------------------------------------------------- leaves
// Leaf sums over a node's positions, exact and independent of the order of
// the rows: each row's g and h become 62-bit fixed-point integers,
// round(g * 2^sg) and round(h * 2^sh), with the shifts taken from the max |g|
// and max |h| of the leaf's tree (its root's statistics), and the integers
// are summed in 128 bits.  Each workgroup of 256 threads takes kLeafChunk consecutive
// positions; the node adds its chunk sums, converts each total to float64
// once and forms f32(G / (H + lambda)).  A row's rounding error is at most
// 2^-63 of the tree's largest |g| (|h|).
constexpr int kLeafChunk = kPartChunk;
constexpr int kLeafBits = 62;

struct LeafShift {
  int sg, sh;
};

// The scale of leaf j's sums: the statistics of its tree (class j /
// nodes_per_class), the same for every leaf of the tree in both engines.
__device__ __forceinline__ LeafShift leaf_shift(const LeafArgs& a, int j) {
  const int32_t* st = a.stats + (int64_t)(j / a.nodes_per_class) * kStatWords;
  return LeafShift{fixed_shift(__builtin_bit_cast(float, st[0]), kLeafBits),
                   fixed_shift(__builtin_bit_cast(float, st[3]), kLeafBits)};
}

__device__ __forceinline__ void leaf_gh(const LeafArgs& a, int p, double& g, double& h) {
  if (a.gh) {
    const float2 v = a.gh[p];
    g = v.x; h = v.y;
  } else {
    const int r = a.perm[p];
    g = a.g[r]; h = a.h[r];
  }
}

// Exact fixed-point sums of g and h over positions [c0, c1) of leaf j into
// part[2 * chunk ..] (any workgroup of whole waves, at most 1024 threads).
__device__ __forceinline__ void leaf_partial_rows(const LeafArgs& a, int chunk, int j, int c0, int c1) {
  const LeafShift ls = leaf_shift(a, j);
  I128 sg, sh;
  for (int p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
    double g, h;
    leaf_gh(a, p, g, h);
    sg.add(to_fixed((float)g, ls.sg));
    sh.add(to_fixed((float)h, ls.sh));
  }
  // per wave (shuffles), then the waves' sums in LDS
  sg = wave_sum128(sg);
  sh = wave_sum128(sh);
  __shared__ unsigned long long s_lo[2][16];
  __shared__ long long s_hi[2][16];
  const int t = threadIdx.x;
  __syncthreads();  // the arrays may still be read by a previous call
  if ((t & 63) == 0) {
    s_lo[0][t >> 6] = sg.lo; s_hi[0][t >> 6] = sg.hi;
    s_lo[1][t >> 6] = sh.lo; s_hi[1][t >> 6] = sh.hi;
  }
  __syncthreads();
  if (t < 2) {
    I128 x;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      I128 y;
      y.lo = s_lo[t][w]; y.hi = s_hi[t][w];
      x.add(y);
    }
    a.part[2 * (int64_t)chunk + t] = x;
  }
}

__device__ __forceinline__ void leaf_partial_block(const LeafArgs& a, int chunk) {
  const int nseg = *a.nseg_dev;
  if (chunk >= a.seg_first[nseg]) return;
  const int j = upper_index(a.seg_first, nseg, chunk);
  const int k = chunk - a.seg_first[j];
  const int c0 = a.node_start[j] + k * kLeafChunk;
  const int c1 = min(c0 + kLeafChunk, a.node_start[j] + a.node_count[j]);
  leaf_partial_rows(a, chunk, j, c0, c1);
}

__global__ void __launch_bounds__(256)
k_leaf_partial(LeafArgs a) { leaf_partial_block(a, blockIdx.x); }

// Leaf j's totals from its chunk sums (exact: any order gives the same
// integers), each converted to float64 once, and its value
// f32(G / (H + lambda)).  The chunk sums are loaded U at a time (one round
// trip per U chunks).
__device__ __forceinline__ float leaf_value(const LeafArgs& a, int j, double& sg, double& sh) {
  const int first = a.seg_first[j], last = a.seg_first[j + 1];
  I128 tg, th;
  constexpr int U = 8;
  for (int c0 = first; c0 < last; c0 += U) {
    I128 vg[U], vh[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t c = min(c0 + u, last - 1);
      vg[u] = a.part[2 * c];
      vh[u] = a.part[2 * c + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c0 + u < last) {
        tg.add(vg[u]);
        th.add(vh[u]);
      }
    }
  }
  const LeafShift ls = leaf_shift(a, j);
  sg = ldexp(tg.to_double(), -ls.sg);
  sh = ldexp(th.to_double(), -ls.sh);
  return (float)(sg / (sh + a.lam));
}

// Chunk `chunk` of leaf j: the leaf value, written once per node (chunk 0),
// and pred[row*K + class] += f32(lr) * leaf for the chunk's rows.
__device__ __forceinline__ void leaf_apply_node(const LeafArgs& a, int chunk, int j) {
  const int k = chunk - a.seg_first[j];
  __shared__ float s_v;
  if (threadIdx.x == 0) {
    double sg, sh;
    const float v = leaf_value(a, j, sg, sh);
    s_v = v;
    if (k == 0) {
      a.leaf_val[j] = v;
      if (a.node_sum) { a.node_sum[2 * j] = sg; a.node_sum[2 * j + 1] = sh; }
    }
  }
  __syncthreads();
  if (a.pred == nullptr) return;
  const int c0 = a.node_start[j] + k * kLeafChunk;
  const int c1 = min(c0 + kLeafChunk, a.node_start[j] + a.node_count[j]);
  const float step = a.lr * s_v;
  const int cls = j / a.nodes_per_class;
  for (int p = c0 + threadIdx.x; p < c1; p += 256) {
    const int r = a.perm[p];
    if ((uint64_t)r >= (uint64_t)a.n_rows) continue;
    float* d = a.pred + (int64_t)r * a.pred_stride + cls;
    *d = *d + step;
  }
}

__device__ __forceinline__ void leaf_apply_block(const LeafArgs& a, int chunk) {
  const int nseg = *a.nseg_dev;
  if (chunk >= a.seg_first[nseg]) return;
  leaf_apply_node(a, chunk, upper_index(a.seg_first, nseg, chunk));
}

__global__ void __launch_bounds__(256)
k_leaf_apply(LeafArgs a) { leaf_apply_block(a, blockIdx.x); }

PlanJob leaf_plan_job(const LeafArgs& a) {
  PlanJob j{};
  j.src.mode = kPlanLeaves;
  j.src.nseg = a.n_nodes;
  j.src.nseg_out = a.nseg_dev;
  j.src.node_active = a.node_active;
  j.src.dec_feat = a.dec_feat;
  j.src.node_count = a.node_count;
  j.chunk_rows = kLeafChunk;
  j.seg_first = a.seg_first;
  j.multi = a.multi;
  j.n_multi = a.n_multi;
  return j;
}

int leaf_launch(const LeafArgs& a, int64_t rows_max, hipStream_t st, bool planned) {
  if (a.n_nodes == 0) return kOk;
  WGBM_REQUIRE(a.node_active && a.node_start && a.node_count && a.seg_first && a.nseg_dev && a.multi && a.n_multi &&
                   a.perm && (a.gh || (a.g && a.h)) && a.leaf_val && a.part && a.stats,
               "leaf: missing table or scratch pointer");
  if (!planned) {
    const PlanJob j = leaf_plan_job(a);
    plan_launch(j.src, j.chunk_rows, j.seg_first, j.multi, j.n_multi, st);
    WGBM_LAUNCHED("k_hist_plan(leaf)");
  }
  const int64_t max_chunks = ceil_div(rows_max, kLeafChunk) + a.n_nodes;
  hipLaunchKernelGGL(k_leaf_partial, dim3((unsigned)max_chunks), dim3(256), 0, st, a);
  WGBM_LAUNCHED("k_leaf_partial");
  hipLaunchKernelGGL(k_leaf_apply, dim3((unsigned)max_chunks), dim3(256), 0, st, a);
  WGBM_LAUNCHED("k_leaf_apply");
  return kOk;
}

int64_t leaf_max_chunks(int64_t rows_max, int64_t n_nodes) { return ceil_div(rows_max, kLeafChunk) + n_nodes; }

// ----------------------------------------------------------------- fused level tail
// Tree record of one level, thread per (node, era): era 0's thread writes
// the node's fields at heap index (2^d - 1) + node of class c's tree of round
// round_ctr[0]; a leaf's value comes from its chunk sums exactly as
// leaf_apply_block forms it.
__device__ __forceinline__ void record_thread(const RecordArgs& r, const LeafArgs& la, int64_t t) {
  if (t >= (int64_t)r.n_nodes * r.E) return;
  const int j = (int)(t / r.E);
  const int e0 = (int)(t - (int64_t)j * r.E);
  const int NH = (2 << r.D) - 1;
  const int c = j >> r.d, jj = j & ((1 << r.d) - 1);
  const int64_t o = ((int64_t)r.round_ctr[0] * r.K + c) * NH + ((1 << r.d) - 1) + jj;
  int st = 0, f = -1, b = -1;
  float v = 0.0f;
  if (r.active[j]) {
    if (r.d < r.D && r.dec_feat[j] >= 0) {
      st = 1;
      f = r.feat_idx ? r.feat_idx[r.dec_feat[j]] : r.dec_feat[j];
      b = r.dec_bin[j];
    } else {
      st = 2;
      double sg, sh;
      v = leaf_value(la, j, sg, sh);
    }
  }
  if (e0 == 0) {
    r.status[o] = st;
    r.feat[o] = f;
    r.bin[o] = b;
    r.value[o] = v;
    r.count[o] = r.active[j] ? r.node_count[j] : 0;
  }
  if (r.out_era_gain) r.out_era_gain[o * r.E + e0] = (st == 1) ? r.era_gain[(int64_t)j * r.E + e0] : 0.0f;
}

// Partition counts of the split segments (workgroups [0, n_count)) and leaf
// chunk sums (the rest).
__global__ void __launch_bounds__(256)
k_count_leafsum(PartArgs p, LeafArgs la, int n_count) {
  const int b = blockIdx.x;
  if (b < n_count) part_count_block(p, b);
  else leaf_partial_block(la, b - n_count);
}

// Stable scatter of the split segments (workgroups [0, n_move)), leaf score
// updates ([n_move, n_move + n_apply)) and the level's tree record (the
// rest, 256 (node, era) pairs per workgroup).
__global__ void __launch_bounds__(256)
k_move_apply_record(PartArgs p, LeafArgs la, RecordArgs r, int n_move, int n_apply) {
  const int b = blockIdx.x;
  if (b < n_move) part_scatter_block(p, b);
  else if (b < n_move + n_apply) leaf_apply_block(la, b - n_move);
  else record_thread(r, la, (int64_t)(b - n_move - n_apply) * 256 + threadIdx.x);
}

int level_tail_launch(const PartArgs* p, const LeafArgs& la, const RecordArgs& r, int64_t rows_max, hipStream_t st,
                      const PlanJob* next) {
  const int nseg = p ? p->n_nodes * p->n_eras : 0;
  if (p) {
    WGBM_REQUIRE(p->codes && p->node_active && p->dec_feat && p->dec_bin && p->seg_start && p->seg_count &&
                     p->node_start && p->node_count && p->perm_in && p->perm_out && p->flag && p->seg_first &&
                     p->nseg_dev && p->lbase && p->rbase && p->cscan && p->segL && p->segP && p->c_seg_start &&
                     p->c_seg_count && p->c_active && p->c_node_start && p->c_node_count && p->b_seg_start &&
                     p->b_seg_count && p->b_seg_slot && p->n_build && p->sib && p->n_split && p->chunk_rows > 0 &&
                     p->c_stats && p->stat_part,
                 "level tail: missing partition table or scratch pointer");
  }
  WGBM_REQUIRE(la.node_active && la.node_start && la.node_count && la.seg_first && la.nseg_dev && la.perm &&
                   (la.gh || (la.g && la.h)) && la.leaf_val && la.part && r.active && r.node_count && r.round_ctr &&
                   r.status && r.feat && r.bin && r.value && r.count && (r.d == r.D || (r.dec_feat && r.dec_bin)),
               "level tail: missing leaf or record pointer");
  const int64_t n_count = p ? ceil_div(rows_max, p->chunk_rows) + nseg : 0;
  const int64_t n_leaf = leaf_max_chunks(rows_max, la.n_nodes);
  const int64_t n_rec = ceil_div((int64_t)r.n_nodes * r.E, 256);
  if (n_count + n_leaf + n_rec > 0x7fffffff) { set_error("level tail grid too large"); return kTooBig; }
  hipLaunchKernelGGL(k_count_leafsum, dim3((unsigned)(n_count + n_leaf)), dim3(256), 0, st,
                     p ? *p : PartArgs{}, la, (int)n_count);
  WGBM_LAUNCHED("k_count_leafsum");
  if (p) {
    const int rc = part_scan_launch(*p, n_count, st);
    if (rc != kOk) return rc;
  }
  hipLaunchKernelGGL(k_move_apply_record, dim3((unsigned)(n_count + n_leaf + n_rec)), dim3(256), 0, st,
                     p ? *p : PartArgs{}, la, r, (int)n_count, (int)n_leaf);
  WGBM_LAUNCHED("k_move_apply_record");
  if (p) {
    if (next) {
      hipLaunchKernelGGL(k_child_stats_plan, dim3((unsigned)(2 * p->n_nodes + 1)), dim3(kPlanThreads), 0, st, *p, *next);
      WGBM_LAUNCHED("k_child_stats_plan");
    } else {
      hipLaunchKernelGGL(k_child_stats, dim3((unsigned)(2 * p->n_nodes)), dim3(256), 0, st, *p);
      WGBM_LAUNCHED("k_child_stats");
    }
  }
  return kOk;
}

// ----------------------------------------------------------------- one-era level tail, decisions folded in
// A one-era level's chunks are mapped to nodes before anything is known
// about the decisions: active node j owns max(1, ceil(count_j / kTailChunk))
// chunks, in node order (a node without rows owns one empty chunk, so that
// its decision and record are written).  Every workgroup of the first launch
// computes that map itself (n_nodes <= kTailMaxNodes), reduces its node's
// scan candidates to the decision (the same reduction pick_node runs) and
// then counts go-left rows (split) or sums the leaf's rows (leaf): the
// separate decision launch and the leaf and partition plans are gone.
// Leaves use the partition's chunk, so their sums are formed exactly as the
// leaf plan forms them.
constexpr int kTailMaxNodes = kScanLdsSeg;
static_assert(kTailMaxNodes == kScanLdsSegMax, "fused levels: the cursor arrays hold kScanLdsSegMax nodes");

// Rows per chunk of a fused level and its workgroup: 4096 rows by 1024
// threads (four passes), so that a node's chunks -- each claiming its place
// with one atomic on the node's cursor, serialised at the memory -- are
// fewer (the root of a million rows: 245 claims, not 977).
constexpr int kTailChunk = 4096;
constexpr int kTailThreads = 1024;

bool level_tail_fusable(int64_t n_nodes, int n_eras, int64_t rows_max) {
  return n_eras == 1 && n_nodes >= 1 && n_nodes <= kTailMaxNodes && rows_max < 0x7fffffff;
}

// The chunk map (every thread of the workgroup calls it): returns
// the node owning `chunk` (-1 past the last chunk) and its first chunk.
// Workgroup 0 stores the first-chunk table and the node count for the later
// launches of the level.
__device__ __forceinline__ int tail_map(const LeafArgs& la, const TailMap& m, int chunk, int& first) {
  __shared__ int s_first[kTailMaxNodes + 1];
  __shared__ int s_node;
  const int n = la.n_nodes, t = threadIdx.x;
  auto chunks_of = [&](int j) -> int {
    if (j >= n || !la.node_active[j]) return 0;
    const int c = la.node_count[j];
    return c <= 0 ? 1 : (c + kTailChunk - 1) / kTailChunk;
  };
  const int j0 = 2 * t, j1 = 2 * t + 1;
  const int n0 = chunks_of(j0), n1 = chunks_of(j1);
  if (t == 0) s_node = -1;
  int tot;
  const int pre = block_exscan(n0 + n1, &tot);  // its barriers order s_node's reset before the writes below
  if (j0 < n) s_first[j0] = pre;
  if (j1 < n) s_first[j1] = pre + n0;
  if (t == 0) s_first[n] = tot;
  if (n0 > 0 && chunk >= pre && chunk < pre + n0) s_node = j0;
  if (n1 > 0 && chunk >= pre + n0 && chunk < pre + n0 + n1) s_node = j1;
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int j = t; j <= n; j += blockDim.x) m.seg_first[j] = s_first[j];
    if (t == 0) *m.nseg = n;
  }
  const int j = s_node;
  first = j >= 0 ? s_first[j] : 0;
  return j;
}

// A chunk's rows in four passes of kTailThreads positions: go-left flags,
// block counts and in-block ranks.
struct ChunkRows {
  static constexpr int U = kTailChunk / kTailThreads;
  int row[U];
  float2 gh[U];
  bool left[U], ok[U];
  int lrank[U], rrank[U];  // rank among the chunk's left (right) rows, in position order
  int L, R;                // the chunk's left and right rows
};

// The chunk's row ids and (g, h): issued before the node's decision is
// known, so that they are in flight meanwhile.
__device__ __forceinline__ void chunk_load(const PartArgs& a, int c0, int c1, ChunkRows& cr) {
#pragma unroll
  for (int u = 0; u < ChunkRows::U; ++u) {
    const int p = c0 + u * kTailThreads + (int)threadIdx.x;
    cr.ok[u] = p < c1;
    const int q = max(min(p, c1 - 1), 0);  // always a valid position (an empty chunk may start at 0)
    cr.row[u] = a.perm_in[q];
    cr.gh[u] = a.gh_in[q];
  }
}

// Go-left flags, counts and ranks of the loaded rows (split on column col at bin).
__device__ __forceinline__ void chunk_rows(const PartArgs& a, int col, int bin, ChunkRows& cr) {
  constexpr int U = ChunkRows::U;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < U; ++u) cr.left[u] = cr.ok[u] && goes_left(a, cr.row[u], col, bin);
  constexpr int NW = kTailThreads / kWave;
  static_assert(U * NW <= kWave, "one wave scans the (pass, wave) counts");
  __shared__ int s_n[U * NW][2];  // left and right rows of (pass, wave), then their exclusive prefixes
  __shared__ int s_tot[2];
  uint64_t lm[U], rm[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    lm[u] = __ballot(cr.left[u]);
    rm[u] = __ballot(cr.ok[u] && !cr.left[u]);
    if (lane == 0) { s_n[u * NW + wv][0] = __popcll(lm[u]); s_n[u * NW + wv][1] = __popcll(rm[u]); }
  }
  __syncthreads();
  if (wv == 0) {  // exclusive scan over (pass, wave) in position order
    int xl = lane < U * NW ? s_n[lane][0] : 0, xr = lane < U * NW ? s_n[lane][1] : 0;
    const int vl = xl, vr = xr;
    for (int o = 1; o < kWave; o <<= 1) {
      const int yl = __shfl_up(xl, o), yr = __shfl_up(xr, o);
      if (lane >= o) { xl += yl; xr += yr; }
    }
    if (lane < U * NW) { s_n[lane][0] = xl - vl; s_n[lane][1] = xr - vr; }
    if (lane == kWave - 1) { s_tot[0] = xl; s_tot[1] = xr; }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    cr.lrank[u] = s_n[u * NW + wv][0] + lanes_below(lm[u]);
    cr.rrank[u] = s_n[u * NW + wv][1] + lanes_below(rm[u]);
  }
  const int lb = s_tot[0], rb = s_tot[1];
  cr.L = lb;
  cr.R = rb;
}

// Node j's cursor: one 64-bit word, the left rows claimed in its high half and
// the right rows in its low half, alone in its 256-byte line (the claims on
// different nodes go to different cache lines and channels).
__device__ __forceinline__ unsigned long long tail_claim(const TailCursors& tc, int j, int L, int R) {
  return atomicAdd(&tc.cur[(int64_t)j * kCurStride], ((unsigned long long)(uint32_t)L << 32) | (uint32_t)R);
}

__device__ __forceinline__ int tail_count(const unsigned long long* cur, int j, int side) {
  const unsigned long long v = cur[(int64_t)j * kCurStride];
  return side ? (int)(uint32_t)v : (int)(v >> 32);
}

// Single-pass split of one chunk of node j: the chunk claims its place in
// each child with one atomic add on the node's cursor (left rows fill the
// node's range from its start, right rows from its end), then writes its
// rows there in position order.  Which chunk comes first in a
// child depends on scheduling; nothing downstream depends on the order of a
// node's rows (histograms and leaf sums are exact integer sums).
__device__ __forceinline__ void chunk_scatter(const PartArgs& a, const TailCursors& tc, int chunk, int j, int col,
                                              int bin, ChunkRows& cr) {
  chunk_rows(a, col, bin, cr);
  __shared__ int s_base[2];
  if (threadIdx.x == 0) {
    const unsigned long long old = tail_claim(tc, j, cr.L, cr.R);
    s_base[0] = (int)(old >> 32);
    s_base[1] = (int)(uint32_t)old;
  }
  __syncthreads();
  const int lo = a.node_start[j] + s_base[0];
  const int hi = a.node_start[j] + a.node_count[j] - s_base[1] - cr.R;  // this chunk's right rows start here
  StatAcc sl, sr;
#pragma unroll
  for (int u = 0; u < ChunkRows::U; ++u) {
    if (!cr.ok[u]) continue;
    const int dst = cr.left[u] ? lo + cr.lrank[u] : hi + cr.rrank[u];
    a.perm_out[dst] = cr.row[u];
    a.gh_out[dst] = cr.gh[u];
    if (cr.left[u]) sl.add(cr.gh[u].x, cr.gh[u].y);
    else sr.add(cr.gh[u].x, cr.gh[u].y);
  }
  sl.block_reduce();
  sr.block_reduce();
  if (threadIdx.x == 0) {  // one partial per chunk and side, folded per child by k_tail_next
    sl.store(a.stat_part + ((int64_t)chunk * 2 + 0) * kStatWords);
    sr.store(a.stat_part + ((int64_t)chunk * 2 + 1) * kStatWords);
  }
}

// A chunk of a node whose children are leaves (the last split level): no
// rows move; the chunk's exact sums of g and h per side (the children's
// leaf sums, scaled as leaf_partial_rows scales them) and its row counts.
__device__ __forceinline__ void chunk_child_sums(const PartArgs& a, const LeafArgs& la, const TailCursors& tc,
                                                 int chunk, int j, int col, int bin, ChunkRows& cr) {
  chunk_rows(a, col, bin, cr);
  if (tc.row_leaf) {  // heap children 2h + 1 (left) and 2h + 2 (right)
    const int h = 2 * (tc.node_base + j) + 1;
#pragma unroll
    for (int u = 0; u < ChunkRows::U; ++u)
      if (cr.ok[u]) tc.row_leaf[cr.row[u]] = (int16_t)(h + (cr.left[u] ? 0 : 1));
  }
  const LeafShift ls = leaf_shift(la, j);  // children belong to j's tree
  I128 acc[4];  // left g, left h, right g, right h
#pragma unroll
  for (int u = 0; u < ChunkRows::U; ++u) {
    if (!cr.ok[u]) continue;
    const int side = cr.left[u] ? 0 : 2;
    acc[side].add(to_fixed(cr.gh[u].x, ls.sg));
    acc[side + 1].add(to_fixed(cr.gh[u].y, ls.sh));
  }
  // per wave (shuffles), then the waves' sums in LDS
  constexpr int NW = kTailThreads / kWave;
  __shared__ unsigned long long s_lo[4][NW];
  __shared__ long long s_hi[4][NW];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const I128 w = wave_sum128(acc[q]);
    if ((t & 63) == 0) { s_lo[q][t >> 6] = w.lo; s_hi[q][t >> 6] = w.hi; }
  }
  __syncthreads();
  if (t < 4) {
    I128 x;
    for (int w = 0; w < NW; ++w) {
      I128 y;
      y.lo = s_lo[t][w]; y.hi = s_hi[t][w];
      x.add(y);
    }
    tc.child_part[4 * (int64_t)chunk + t] = x;
  }
  if (t == 0) tail_claim(tc, j, cr.L, cr.R);  // the children's row counts
}

// The level's one partition launch: decisions (as k_level_pick forms them),
// then per chunk a single-pass split, the children's leaf sums (last split
// level) or, for a leaf, its chunk sums.
__global__ void __launch_bounds__(kTailThreads)
k_tail_part(PartArgs p, LeafArgs la, TailPick pk, TailMap m, TailCursors tc, int last) {
  const int chunk = blockIdx.x;
  int first;
  const int j = tail_map(la, m, chunk, first);
  if (j < 0) return;
  const int k = chunk - first;
  if (threadIdx.x == 0) m.chunk_node[chunk] = j;
  const int c0 = la.node_start[j] + k * kTailChunk;
  const int c1 = min(c0 + kTailChunk, la.node_start[j] + max(la.node_count[j], 0));
  ChunkRows cr;
  chunk_load(p, c0, c1, cr);
  __shared__ int s_fb[2];
  if (threadIdx.x < kWave) {
    const Cand b = node_best(pk.cand_agree, pk.cand_crit, pk.cand_idx, j, pk.wg_per_node, threadIdx.x);
    int f_, b_;
    cand_split(b, pk.n_bins, f_, b_);
    if (threadIdx.x == 0) {
      s_fb[0] = f_;
      s_fb[1] = b_;
      if (k == 0) {  // one workgroup per node writes the decision
        pk.dec_feat[j] = f_;
        pk.dec_bin[j] = b_;
        if (pk.era_gain) pk.era_gain[j] = f_ >= 0 ? b.crit : 0.0f;  // one era: the winner's gain
      }
    }
  }
  __syncthreads();
  const int fl = s_fb[0], bin = s_fb[1];
  if (fl < 0) {
    if (tc.row_leaf) {
#pragma unroll
      for (int u = 0; u < ChunkRows::U; ++u)
        if (cr.ok[u]) tc.row_leaf[cr.row[u]] = (int16_t)(tc.node_base + j);
    }
    leaf_partial_rows(la, chunk, j, c0, c1);
    return;
  }
  const int col = p.feat_idx ? p.feat_idx[fl] : fl;
  if (last) chunk_child_sums(p, la, tc, chunk, j, col, bin, cr);
  else chunk_scatter(p, tc, chunk, j, col, bin, cr);
}

// Tree record of a child of the last split level (child c of node c / 2),
// one wave per child: a leaf whose sums come from its parent's chunks (the
// lanes take the chunks, the exact sums add in any order), or nothing.
__device__ __forceinline__ void record_child(const RecordArgs& r, const PartArgs& p, const LeafArgs& la,
                                             const TailMap& m, const TailCursors& tc, int c, int lane) {
  const int j = c >> 1, side = c & 1;
  const int NH = (2 << r.D) - 1;
  const int d = r.d + 1;  // the children's level
  const int cls = c >> d, cc = c & ((1 << d) - 1);
  const int64_t o = ((int64_t)r.round_ctr[0] * r.K + cls) * NH + ((1 << d) - 1) + cc;
  const bool leaf = p.node_active[j] && p.dec_feat[j] >= 0;
  float v = 0.0f;
  if (leaf) {
    I128 tg, th;
    for (int q = m.seg_first[j] + lane; q < m.seg_first[j + 1]; q += kWave) {
      tg.add(tc.child_part[4 * (int64_t)q + 2 * side]);
      th.add(tc.child_part[4 * (int64_t)q + 2 * side + 1]);
    }
    tg = wave_sum128(tg);
    th = wave_sum128(th);
    const LeafShift ls = leaf_shift(la, j);
    const double sg = ldexp(tg.to_double(), -ls.sg), sh = ldexp(th.to_double(), -ls.sh);
    v = (float)(sg / (sh + la.lam));
  }
  if (lane != 0) return;
  r.status[o] = leaf ? 2 : 0;
  r.feat[o] = -1;
  r.bin[o] = -1;
  r.value[o] = v;
  r.count[o] = leaf ? tail_count(tc.cur, j, side) : 0;
  if (r.out_era_gain) r.out_era_gain[o] = 0.0f;
}

// After the partition launch: workgroup 0 builds the child tables from the
// cursors (left child at the node's start, right child after it), the
// smaller-child build list (left on ties: warpgbm/core.py:1438) and sibling
// map, clears the next level's cursors and runs the next level's histogram
// plan; workgroups 1 .. 2n fold the children's statistics; the rest write
// the level's tree record (and, after the last split level, the leaves').
__global__ void __launch_bounds__(kPlanThreads)
k_tail_next(PartArgs p, LeafArgs la, RecordArgs r, TailMap m, TailCursors tc, const PlanJob next, int has_plan,
            int last) {
  const int n = p.n_nodes, t = threadIdx.x;
  const int b = blockIdx.x;
  const int n_stat = last ? 0 : 2 * n;
  if (b == 0) {
    if (last) return;
    __shared__ int s_cnt[2 * kTailMaxNodes];
    for (int j = t; j < n; j += blockDim.x) {
      const int l = 2 * j, rr = 2 * j + 1;
      const bool sp = p.node_active[j] && p.dec_feat[j] >= 0;
      const int start = p.node_start[j];
      const int Lt = sp ? tail_count(tc.cur, j, 0) : 0, Rt = sp ? tail_count(tc.cur, j, 1) : 0;
      p.c_active[l] = sp; p.c_active[rr] = sp;
      p.c_node_start[l] = sp ? start : 0; p.c_node_count[l] = Lt;
      p.c_node_start[rr] = sp ? start + Lt : 0; p.c_node_count[rr] = Rt;
      p.c_seg_start[l] = p.c_node_start[l]; p.c_seg_count[l] = Lt;  // one era: a node is its segment
      p.c_seg_start[rr] = p.c_node_start[rr]; p.c_seg_count[rr] = Rt;
      s_cnt[l] = sp ? Lt : -1;
      s_cnt[rr] = Rt;
    }
    // the next level's cursors (it uses them only if it is fused: at most kTailMaxNodes nodes)
    for (int q = t; q < min(2 * n, kTailMaxNodes); q += blockDim.x) tc.cur_next[(int64_t)q * kCurStride] = 0ull;
    __syncthreads();
    const int sp = t < n && s_cnt[2 * t] >= 0;
    int ns;
    const int rank = block_exscan(sp, &ns);
    if (sp) {
      const int l = 2 * t, rr = 2 * t + 1;
      const int built = s_cnt[l] <= s_cnt[rr] ? l : rr;
      p.sib[3 * rank + 0] = t;
      p.sib[3 * rank + 1] = built;
      p.sib[3 * rank + 2] = built ^ 1;
      p.b_seg_start[rank] = built == l ? p.node_start[t] : p.node_start[t] + s_cnt[l];
      p.b_seg_count[rank] = s_cnt[built];
      p.b_seg_slot[rank] = built;
    }
    if (t == 0) {
      *p.n_split = ns;
      *p.n_build = ns;
    }
    __syncthreads();  // the build list is complete (and visible to this workgroup) before its plan
    if (has_plan) plan_body(next.src, next.chunk_rows, next.seg_first, next.seg_pfirst, next.multi, next.n_multi);
    return;
  }
  if (b <= n_stat) {
    const int c = b - 1;
    const int j = c >> 1, side = c & 1;
    StatAcc acc;
    if (p.node_active[j] && p.dec_feat[j] >= 0)
      for (int q = m.seg_first[j] + t; q < m.seg_first[j + 1]; q += blockDim.x)
        acc.load_merge(p.stat_part + ((int64_t)q * 2 + side) * kStatWords);
    acc.block_reduce();
    if (t == 0) acc.store(p.c_stats + (int64_t)c * kStatWords);
    return;
  }
  const int n_rec = (n + blockDim.x - 1) / blockDim.x;  // workgroups of the level's own records
  const int rb = b - 1 - n_stat;
  if (rb < n_rec) {
    const int q = rb * blockDim.x + t;
    if (q < n) record_thread(r, la, q);
  } else if (last) {  // a wave per child
    const int c = (rb - n_rec) * (blockDim.x / kWave) + (t >> 6);
    if (c < 2 * n) record_child(r, p, la, m, tc, c, t & 63);
  }
}

int level_fused_launch(const PartArgs& p, const LeafArgs& la, const RecordArgs& r, const TailPick& pick,
                       const TailMap& m, const TailCursors& tc, bool last, int64_t rows_max, hipStream_t st,
                       const PlanJob* next) {
  WGBM_REQUIRE(level_tail_fusable(la.n_nodes, r.E, rows_max) && p.n_eras == 1 && p.n_nodes == la.n_nodes &&
                   r.n_nodes == la.n_nodes && r.d < r.D,
               "level tail: not a one-era level of at most %d nodes before the last", kTailMaxNodes);
  WGBM_REQUIRE(m.seg_first && m.nseg && m.chunk_node && la.seg_first == m.seg_first && la.nseg_dev == m.nseg &&
                   p.chunk_seg == m.chunk_node && tc.cur && tc.cur_next && (!last || tc.child_part),
               "level tail: the leaf tables must be the chunk map's");
  WGBM_REQUIRE(pick.cand_agree && pick.cand_crit && pick.cand_idx && pick.wg_per_node >= 1 &&
                   pick.dec_feat == p.dec_feat && pick.dec_bin == p.dec_bin && r.dec_feat == p.dec_feat &&
                   r.dec_bin == p.dec_bin && la.dec_feat == p.dec_feat,
               "level tail: the decisions must go where the partition and the record read them");
  WGBM_REQUIRE(la.node_active && la.node_start && la.node_count && la.perm && la.gh && la.stats && la.part &&
                   la.pred == nullptr && p.codes && p.perm_in && p.gh_in && p.perm_out && p.gh_out &&
                   p.c_seg_start && p.c_seg_count && p.c_active && p.c_node_start && p.c_node_count &&
                   p.b_seg_start && p.b_seg_count && p.b_seg_slot && p.n_build && p.sib && p.n_split &&
                   p.c_stats && p.stat_part && r.active && r.node_count && r.round_ctr && r.status && r.feat &&
                   r.bin && r.value && r.count,
               "level tail: missing table or scratch pointer");
  WGBM_REQUIRE(last || next, "level tail: the next level needs its histogram plan");
  WGBM_REQUIRE(!tc.row_leaf || (la.n_nodes == la.nodes_per_class && tc.node_base == la.n_nodes - 1 &&
                                4 * (int64_t)la.n_nodes - 2 <= 32767),
               "level tail: row leaves are recorded for one class, heap indices below 2^15");
  const int64_t n_chunks = ceil_div(rows_max, kTailChunk) + la.n_nodes;
  hipLaunchKernelGGL(k_tail_part, dim3((unsigned)n_chunks), dim3(kTailThreads), 0, st, p, la, pick, m, tc,
                     last ? 1 : 0);
  WGBM_LAUNCHED("k_tail_part");
  const int64_t n_stat = last ? 0 : 2 * (int64_t)la.n_nodes;
  const int64_t n_rec = ceil_div((int64_t)la.n_nodes, kPlanThreads) +
                        (last ? ceil_div(2 * (int64_t)la.n_nodes, kPlanThreads / kWave) : 0);
  hipLaunchKernelGGL(k_tail_next, dim3((unsigned)(1 + n_stat + n_rec)), dim3(kPlanThreads), 0, st, p, la, r, m, tc,
                     next ? *next : PlanJob{}, next ? 1 : 0, last ? 1 : 0);
  WGBM_LAUNCHED("k_tail_next");
  return kOk;
}

// ----------------------------------------------------------------- subtract
// other = parent - built for every split of the previous level
// (warpgbm/core.py:1445-1455), float32 elementwise over [E][F][B].
__global__ void __launch_bounds__(256)
k_hist_subtract(const float* __restrict__ parent_g, const float* __restrict__ parent_h,
                float* __restrict__ child_g, float* __restrict__ child_h, const int32_t* __restrict__ sib,
                const int32_t* __restrict__ n_split, int64_t node_elems, int blocks_per_node) {
  const int q = blockIdx.x / blocks_per_node;
  if (q >= *n_split) return;
  const int blk = blockIdx.x - q * blocks_per_node;
  const int64_t pj = sib[3 * q], bj = sib[3 * q + 1], oj = sib[3 * q + 2];
  for (int64_t i = (int64_t)blk * 256 * 4 + threadIdx.x; i < node_elems && i < (int64_t)(blk + 1) * 1024;
       i += 256) {
    child_g[oj * node_elems + i] = parent_g[pj * node_elems + i] - child_g[bj * node_elems + i];
    child_h[oj * node_elems + i] = parent_h[pj * node_elems + i] - child_h[bj * node_elems + i];
  }
}

int subtract_launch(const float* pg, const float* ph, float* cg, float* ch, const int32_t* sib,
                    const int32_t* n_split, int64_t max_split, int64_t node_elems, hipStream_t st) {
  if (max_split == 0 || node_elems == 0) return kOk;
  const int bpn = (int)ceil_div(node_elems, 1024);
  hipLaunchKernelGGL(k_hist_subtract, dim3((unsigned)(max_split * bpn)), dim3(256), 0, st, pg, ph, cg, ch, sib,
                     n_split, node_elems, bpn);
  WGBM_LAUNCHED("k_hist_subtract");
  return kOk;
}

__global__ void k_part_single_init(int32_t* sc, int feature, int bin, int n) {
  if (threadIdx.x == 0) { sc[0] = 1; sc[1] = feature; sc[2] = bin; sc[3] = 0; sc[4] = n; }
}

__global__ void k_leaf_single_init(int32_t* sc, int n) {
  if (threadIdx.x == 0) { sc[0] = 1; sc[1] = 0; sc[2] = n; }
}

__global__ void k_sub_plain(const float* pg, const float* ph, const float* cg, const float* ch, float* sg,
                            float* sh, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sg[i] = pg[i] - cg[i];
  sh[i] = ph[i] - ch[i];
}

// ----------------------------------------------------------------- objectives
// Regression: g = (y - F) * w, h = w (warpgbm/core.py:1545-1547).
// Written per row (g, h) and/or per position of order `perm0` (gh).
__global__ void __launch_bounds__(256)
k_grad_regression(const float* __restrict__ y, const float* __restrict__ pred, const float* __restrict__ w,
                  int64_t n, const int32_t* __restrict__ perm0, float* __restrict__ g, float* __restrict__ h,
                  float2* __restrict__ gh) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t r = perm0 ? perm0[p] : p;
  const float wr = w[r];
  const float gr = (y[r] - pred[r]) * wr;
  if (g) { g[r] = gr; h[r] = wr; }
  if (gh) gh[p] = make_float2(gr, wr);
}

// Multi-output squared error with shared splits, one boosting round
// (warpgbm/core.py:1666-1679): per output m, valid = isfinite(y[r, m]),
// g_m = valid ? (y - pred) * w : 0, h_m = valid ? w : 0, written
// output-major (m*n + r); the split targets are the float32 sums over m in
// output order, split_g[r] = sum_m g_m, split_h[r] = sum_m h_m.
__global__ void __launch_bounds__(256)
k_grad_multi_regression(const float* __restrict__ y, const float* __restrict__ pred, const float* __restrict__ w,
                        int64_t n, int M, float* __restrict__ g, float* __restrict__ h,
                        float* __restrict__ split_g, float* __restrict__ split_h) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float sg = 0.0f, sh = 0.0f;
  for (int m = 0; m < M; ++m) {
    const int64_t i = r * M + m;
    const float yv = y[i];
    const bool valid = isfinite(yv);
    const float wm = valid ? w[i] : 0.0f;
    const float gm = valid ? (yv - pred[i]) * wm : 0.0f;
    g[(int64_t)m * n + r] = gm;
    h[(int64_t)m * n + r] = wm;
    sg += gm;
    sh += wm;
  }
  split_g[r] = sg;
  split_h[r] = sh;
}

// Softmax cross-entropy, one boosting round (warpgbm/core.py:339-367,
// :1755-1756; softmax of warpgbm/metrics.py:12-15): p = softmax(F_row),
// g_k = -(p_k - [y == k]) * w, h_k = max(p_k (1 - p_k), 1e-6) * w.
// Output class-major: index k*n + position.
__global__ void __launch_bounds__(256)
k_grad_softmax(const float* __restrict__ F, const int32_t* __restrict__ y, const float* __restrict__ w,
               int64_t n, int K, const int32_t* __restrict__ perm0, float* __restrict__ g,
               float* __restrict__ h, float2* __restrict__ gh) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t r = perm0 ? perm0[p] : p;
  const float* f = F + r * K;
  float mx = f[0];
  for (int k = 1; k < K; ++k) mx = fmaxf(mx, f[k]);
  float den = 0.0f;
  for (int k = 0; k < K; ++k) den += expf(f[k] - mx);
  const float wr = w[r];
  const int yr = y[r];
  for (int k = 0; k < K; ++k) {
    const float pk = expf(f[k] - mx) / den;
    const float gk = -(pk - (yr == k ? 1.0f : 0.0f)) * wr;
    const float hk = fmaxf(pk * (1.0f - pk), 1e-6f) * wr;
    if (g) { g[(int64_t)k * n + r] = gk; h[(int64_t)k * n + r] = hk; }
    if (gh) gh[(int64_t)k * n + p] = make_float2(gk, hk);
  }
}

}  // namespace wgbm

using namespace wgbm;

// ----------------------------------------------------------------- C ABI
extern "C" size_t wgbm_partition_workspace_bytes(int64_t n_samples) {
  if (n_samples < 0) return 0;
  const int64_t mc = ceil_div(n_samples, kPartChunk) + 1;
  return (size_t)(4 * (16 + 4 * mc + 1) + 64 * 4) + (size_t)n_samples + 1024;
}

// Stable split of one node's row list (plain path): rows with
// code[row, feature] <= bin first, then the others, each in input order.
// out_rows[0:n_left] / out_rows[n_left:n]; n_left written to *n_left_out
// (device int32).
extern "C" int wgbm_partition_rows(const int8_t* bin_indices, int64_t n_total_rows, int64_t n_master_feat,
                                   const int32_t* rows, int64_t n, int32_t feature, int32_t bin,
                                   int32_t* out_rows, int32_t* n_left_out, void* workspace,
                                   size_t workspace_bytes, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n < 0x7fffffff, "row count out of range");
  WGBM_REQUIRE(feature >= 0 && feature < n_master_feat, "split feature %d out of range [0, %lld)", feature,
               (long long)n_master_feat);
  WGBM_REQUIRE(workspace_bytes >= wgbm_partition_workspace_bytes(n), "partition workspace too small");
  const int64_t mc = ceil_div(n, kPartChunk) + 1;
  int32_t* w = static_cast<int32_t*>(workspace);
  // scalars / single-node tables at the front, chunk arrays after.
  int32_t* sc = w;            // 64 ints of scalars and tiny tables
  int32_t* lbase = w + 64;
  int32_t* rbase = lbase + mc;
  int32_t* cscan = rbase + mc;   // mc + 1
  int32_t* multi = cscan + mc + 1;
  PartArgs a{};
  a.codes = bin_indices; a.row_stride = n_master_feat; a.n_rows = n_total_rows;
  a.feat_idx = nullptr;
  a.n_nodes = 1; a.n_eras = 1;
  a.node_active = sc + 0;    // 1
  a.dec_feat = sc + 1;       // feature
  a.dec_bin = sc + 2;        // bin
  a.seg_start = sc + 3;      // 0
  a.seg_count = sc + 4;      // n
  a.node_start = sc + 3;
  a.node_count = sc + 4;
  a.seg_first = sc + 6;      // 2 ints
  a.nseg_dev = sc + 8;
  a.n_multi = sc + 9;
  a.segL = sc + 10;
  a.c_seg_start = sc + 12; a.c_seg_count = sc + 14; a.c_active = sc + 16;
  a.c_node_start = sc + 18; a.c_node_count = sc + 20;
  a.b_seg_start = sc + 22; a.b_seg_count = sc + 23; a.b_seg_slot = sc + 24;
  a.n_build = sc + 25; a.n_split = sc + 26; a.sib = sc + 27;  // 3 ints
  a.segP = sc + 32;                                             // 4 ints
  a.multi = multi;
  a.flag = reinterpret_cast<uint8_t*>(multi + mc);
  a.lbase = lbase; a.rbase = rbase; a.cscan = cscan;
  a.perm_in = rows; a.gh_in = nullptr;
  a.perm_out = out_rows; a.gh_out = nullptr;
  a.chunk_rows = kPartChunk;
  hipLaunchKernelGGL(k_part_single_init, dim3(1), dim3(64), 0, stream, sc, feature, bin, (int)n);
  WGBM_LAUNCHED("k_part_single_init");
  const int rc = partition_launch(a, n, stream);
  if (rc != kOk) return rc;
  WGBM_HIP(hipMemcpyAsync(n_left_out, a.c_node_count, sizeof(int32_t), hipMemcpyDeviceToDevice, stream));
  return kOk;
}

extern "C" size_t wgbm_leaf_workspace_bytes(int64_t n_samples) {
  if (n_samples < 0) return 0;
  const int64_t mc = leaf_max_chunks(n_samples, 1);
  // scalars, chunk sums, plan scratch, statistics partials and the leaf's statistics
  return (size_t)(64 * 4 + sizeof(I128) * 2 * mc + 4 * mc + 4 * kStatWords * (kStatBlocks + 1)) + 1024;
}

// Leaf value of one node (plain path): sum(g)/(sum(h)+lambda) over its rows
// (warpgbm/core.py:1377-1380; the sums exact in fixed point as the batched
// engine forms them, scaled by scale_stats -- the statistics of the leaf's
// tree, as the batched engine scales its leaves -- or when that is NULL by
// the leaf's own), written as float32 to *leaf_out, and
// pred[row*stride + col] += f32(lr) * leaf for every row when pred is set.
extern "C" size_t wgbm_row_stats_workspace_bytes(void) { return 4 * kStatWords * (size_t)kStatBlocks; }

// Statistics of a row list (max |g|, min h, max h, max |h| as order-preserving
// integers) into out_stats (device, 4 ints): a tree's root rows give the
// scale of its leaf sums.
extern "C" int wgbm_row_stats(const float* grads, const float* hess, const int32_t* rows, int64_t n,
                              int64_t n_total_rows, int32_t* out_stats, void* workspace, size_t workspace_bytes,
                              hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n < 0x7fffffff, "row count out of range");
  WGBM_REQUIRE(out_stats && workspace && workspace_bytes >= wgbm_row_stats_workspace_bytes(),
               "row statistics: missing output or workspace too small");
  int32_t* part = static_cast<int32_t*>(workspace);
  const int nsb = stats_blocks(n);
  hipLaunchKernelGGL(k_list_stats, dim3((unsigned)nsb), dim3(256), 0, stream, rows, (int)n, n_total_rows, grads, hess,
                     part);
  WGBM_LAUNCHED("k_list_stats");
  hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, stream, part, nsb, out_stats);
  WGBM_LAUNCHED("k_stats_reduce");
  return kOk;
}

extern "C" int wgbm_leaf_value(const float* grads, const float* hess, const int32_t* rows, int64_t n,
                               int64_t n_total_rows, double lambda, float learning_rate, float* pred,
                               int32_t pred_stride, int32_t pred_col, float* leaf_out, const int32_t* scale_stats,
                               void* workspace, size_t workspace_bytes, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n < 0x7fffffff, "row count out of range");
  WGBM_REQUIRE(workspace_bytes >= wgbm_leaf_workspace_bytes(n), "leaf workspace too small");
  WGBM_REQUIRE(pred == nullptr || (pred_stride >= 1 && pred_col >= 0 && pred_col < pred_stride),
               "bad prediction column");
  int32_t* sc = static_cast<int32_t*>(workspace);
  const int64_t mc = leaf_max_chunks(n, 1);
  LeafArgs a{};
  a.n_nodes = 1;
  a.node_active = sc + 0;
  a.dec_feat = nullptr;
  a.node_start = sc + 1;
  a.node_count = sc + 2;
  a.seg_first = sc + 4;  // 2
  a.nseg_dev = sc + 6;
  a.n_multi = sc + 7;
  a.perm = rows; a.gh = nullptr; a.g = grads; a.h = hess;
  a.n_rows = n_total_rows;
  a.lam = lambda; a.lr = learning_rate;
  a.pred = pred;
  a.pred_stride = pred_stride;
  a.nodes_per_class = 1;
  a.leaf_val = leaf_out;
  a.node_sum = nullptr;
  I128* part = reinterpret_cast<I128*>(reinterpret_cast<char*>(workspace) + 256);
  a.part = part;
  a.multi = reinterpret_cast<int32_t*>(part + 2 * mc);
  int32_t* stat_part = a.multi + mc;
  int32_t* stats = stat_part + (int64_t)kStatWords * kStatBlocks;
  a.stats = scale_stats ? scale_stats : stats;
  if (pred) a.pred = pred + pred_col;
  hipLaunchKernelGGL(k_leaf_single_init, dim3(1), dim3(64), 0, stream, sc, (int)n);
  WGBM_LAUNCHED("k_leaf_single_init");
  if (!scale_stats) {  // scaled by the leaf's own max |g| and max |h|
    const int nsb = stats_blocks(n);
    hipLaunchKernelGGL(k_list_stats, dim3((unsigned)nsb), dim3(256), 0, stream, rows, (int)n, n_total_rows, grads,
                       hess, stat_part);
    WGBM_LAUNCHED("k_list_stats");
    hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, stream, stat_part, nsb, stats);
    WGBM_LAUNCHED("k_stats_reduce");
  }
  return leaf_launch(a, n, stream);
}

extern "C" int wgbm_grad_regression(const float* y, const float* pred, const float* w, int64_t n,
                                    float* g_out, float* h_out, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0, "row count out of range");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_grad_regression, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, y, pred, w, n,
                     (const int32_t*)nullptr, g_out, h_out, (float2*)nullptr);
  WGBM_LAUNCHED("k_grad_regression");
  return kOk;
}

extern "C" int wgbm_grad_multi_regression(const float* y, const float* pred, const float* w, int64_t n,
                                          int32_t n_outputs, float* g_out, float* h_out, float* split_g,
                                          float* split_h, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n_outputs >= 1, "bad multi-output shape");
  WGBM_REQUIRE(n * (int64_t)n_outputs < ((int64_t)1 << 62), "multi-output size out of range");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_grad_multi_regression, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, y, pred, w,
                     n, (int)n_outputs, g_out, h_out, split_g, split_h);
  WGBM_LAUNCHED("k_grad_multi_regression");
  return kOk;
}

extern "C" int wgbm_grad_softmax(const float* scores, const int32_t* labels, const float* w, int64_t n,
                                 int32_t n_classes, float* g_out, float* h_out, hipStream_t stream) {
  WGBM_REQUIRE(n >= 0 && n_classes >= 2, "bad softmax shape");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_grad_softmax, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, scores, labels, w,
                     n, (int)n_classes, (const int32_t*)nullptr, g_out, h_out, (float2*)nullptr);
  WGBM_LAUNCHED("k_grad_softmax");
  return kOk;
}

extern "C" int wgbm_hist_subtract(const float* parent_g, const float* parent_h, const float* child_g,
                                  const float* child_h, float* sib_g, float* sib_h, int64_t n,
                                  hipStream_t stream) {
  WGBM_REQUIRE(n >= 0, "element count out of range");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_sub_plain, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, parent_g, parent_h,
                     child_g, child_h, sib_g, sib_h, n);
  WGBM_LAUNCHED("k_sub_plain");
  return kOk;
}
