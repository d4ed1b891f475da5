// Chunk plans (device side): one workgroup turns per-segment row
// counts into chunk offsets, the list of multi-chunk segments and their
// partial slots.  Shared by the histogram, partition and leaf launches and by
// the fused launches that run a plan beside other work.
#pragma once
#include "common.h"
#include "wgbm_internal.h"

namespace wgbm {

__device__ __forceinline__ int plan_count(const PlanSrc& p, int s) {
  if (p.mode == kPlanSegments) return p.seg_count[s];
  if (p.mode == kPlanLeaves) {
    const bool leaf = p.node_active[s] && (p.dec_feat == nullptr || p.dec_feat[s] < 0);
    return leaf ? p.node_count[s] : -1;
  }
  const int j = s / p.n_eras;  // kPlanSplits: segments of split nodes
  const bool split = p.node_active[j] && p.dec_feat[j] >= 0;
  return split ? p.seg_count[s] : -1;
}

// Chunks of segment s: a segment with rows owns ceil(rows / chunk_rows), an
// empty one owns one (empty) chunk so that its slot is written, a padding or
// non-selected segment (count < 0) owns none.
__device__ __forceinline__ int plan_chunks(const PlanSrc& p, int s, int chunk_rows) {
  const int c = plan_count(p, s);
  return c < 0 ? 0 : (c == 0 ? 1 : (c + chunk_rows - 1) / chunk_rows);
}

constexpr int kPlanCpt = 32;        // segments per thread per pass
constexpr int kPlanOwnChunks = 32;  // a segment with more chunks is written by the whole workgroup
constexpr int kPlanBig = 256;       // such segments listed per pass

// Chunk tables of segment s (chunks first .. first + n - 1, partials from pf
// or -1): chunks k = k0, k0 + step, ... get their segment and descriptor.
__device__ __forceinline__ void plan_write_chunks(const PlanSrc& src, int s, int first, int n, int pf, int k0,
                                                  int step, int chunk_rows) {
  int start = 0, end = 0, slot = 0;
  if (src.chunk_desc) {
    start = src.seg_start[s];
    end = start + max(plan_count(src, s), 0);
    slot = src.seg_slot[s];
  }
  for (int k = k0; k < n; k += step) {
    if (src.chunk_seg) src.chunk_seg[first + k] = s;
    if (src.chunk_desc) {
      const int a0 = start + k * chunk_rows;
      src.chunk_desc[first + k] = make_int4(a0, min(a0 + chunk_rows, end), slot, pf >= 0 ? pf + k : -1);
    }
  }
}

// One workgroup (any multiple of 64 threads up to 1024) computes a plan.
__device__ __forceinline__ void plan_body(const PlanSrc& src, int32_t chunk_rows, int32_t* __restrict__ seg_first,
                          int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi,
                          int32_t* __restrict__ n_multi) {
  __shared__ int32_t s_sum[3][1024 / kWave];
  __shared__ int32_t s_carry[3];
  const int nseg = src.nseg_dev ? *src.nseg_dev : src.nseg;
  if (src.nseg_out && threadIdx.x == 0) *src.nseg_out = nseg;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nt = blockDim.x, nw = nt >> 6;
  if (src.chunk_out) {
    // Rows per chunk for this launch: enough chunks to give every slot a
    // work item (chunk_target of them), within [chunk_min, chunk_rows].  Each
    // non-empty segment may end in a ragged chunk, so the rows are spread over
    // chunk_target minus that many chunks (at least half the target): the
    // launch then has at most chunk_target chunks and no slot takes a second
    // work item after the others have finished.
    __shared__ int s_rows, s_live;
    if (t == 0) { s_rows = 0; s_live = 0; }
    __syncthreads();
    int mine = 0, live = 0;
    for (int s = t; s < nseg; s += blockDim.x) {
      const int c = max(plan_count(src, s), 0);
      mine += c;
      live += c > 0;
    }
    for (int o = 32; o >= 1; o >>= 1) { mine += __shfl_xor(mine, o); live += __shfl_xor(live, o); }
    if (lane == 0) { atomicAdd(&s_rows, mine); atomicAdd(&s_live, live); }
    __syncthreads();
    const int tgt = max(src.chunk_target - s_live, (src.chunk_target + 1) / 2);
    const int want = (s_rows + tgt - 1) / tgt;
    chunk_rows = min(max((want + 63) & ~63, src.chunk_min), chunk_rows);  // whole 64-row blocks
    if (t == 0) *src.chunk_out = chunk_rows;
  }
  __shared__ int s_maxnch;  // most chunks of a multi-chunk segment
  __shared__ int s_nbig;
  __shared__ int s_big[kPlanBig][3];  // {segment, first chunk, first partial or -1}
  if (t < 3) s_carry[t] = 0;
  if (t == 0) s_maxnch = 0;
  __syncthreads();
  const bool chunk_tables = src.chunk_desc || src.chunk_seg;
  // Blocked scan: in each pass thread t takes cpt consecutive segments, so a
  // pass covers up to blockDim * kPlanCpt segments with one workgroup scan.
  for (int base = 0; base < nseg; base += nt * kPlanCpt) {
    const int cpt = min(kPlanCpt, (nseg - base + nt - 1) / nt);
    const int s0 = base + t * cpt;
    const int s1 = min(s0 + cpt, nseg);
    int v[3] = {0, 0, 0};  // chunks, multi-chunk segments, their chunks
    for (int s = s0; s < s1; ++s) {
      const int n = plan_chunks(src, s, chunk_rows);
      v[0] += n;
      v[1] += n > 1;
      v[2] += n > 1 ? n : 0;
    }
    int x[3] = {v[0], v[1], v[2]};
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int y = __shfl_up(x[q], o);
        if (lane >= o) x[q] += y;
      }
    }
    if (lane == 63) { s_sum[0][wv] = x[0]; s_sum[1][wv] = x[1]; s_sum[2][wv] = x[2]; }
    if (t == 0) s_nbig = 0;
    __syncthreads();
    int pre[3] = {0, 0, 0};
    for (int k = 0; k < wv; ++k) { pre[0] += s_sum[0][k]; pre[1] += s_sum[1][k]; pre[2] += s_sum[2][k]; }
    const int c0 = s_carry[0], c1 = s_carry[1], c2 = s_carry[2];
    int r0 = c0 + pre[0] + x[0] - v[0];
    int r1 = c1 + pre[1] + x[1] - v[1];
    int r2 = c2 + pre[2] + x[2] - v[2];
    for (int s = s0; s < s1; ++s) {
      const int n = plan_chunks(src, s, chunk_rows);
      const int pf = n > 1 ? r2 : -1;
      seg_first[s] = r0;
      if (seg_pfirst) seg_pfirst[s] = r2;
      if (n > 1) {
        if (multi) multi[r1] = s;
        if (src.multi_desc) {
          src.multi_desc[r1] = make_int4(r2, n, src.seg_slot[s], 0);
          atomicMax(&s_maxnch, n);
        }
      }
      if (chunk_tables && n > 0) {
        // a segment of few chunks writes them itself; larger ones are
        // written by the whole workgroup below (or here if the list is full)
        int q = kPlanBig;
        if (n > kPlanOwnChunks) q = atomicAdd(&s_nbig, 1);
        if (q < kPlanBig) {
          s_big[q][0] = s; s_big[q][1] = r0; s_big[q][2] = pf;
        } else {
          plan_write_chunks(src, s, r0, n, pf, 0, 1, chunk_rows);
        }
      }
      r0 += n;
      if (n > 1) { r1 += 1; r2 += n; }
    }
    __syncthreads();
    if (chunk_tables) {
      const int nbig = min(s_nbig, kPlanBig);
      for (int q = 0; q < nbig; ++q) {
        const int s = s_big[q][0];
        plan_write_chunks(src, s, s_big[q][1], plan_chunks(src, s, chunk_rows), s_big[q][2], t, nt, chunk_rows);
      }
    }
    __syncthreads();
    if (t == nt - 1) {
      int tot[3] = {0, 0, 0};
      for (int k = 0; k < nw; ++k) { tot[0] += s_sum[0][k]; tot[1] += s_sum[1][k]; tot[2] += s_sum[2][k]; }
      s_carry[0] = c0 + tot[0];
      s_carry[1] = c1 + tot[1];
      s_carry[2] = c2 + tot[2];
    }
    __syncthreads();
  }
  if (t == 0) {
    seg_first[nseg] = s_carry[0];
    if (n_multi) *n_multi = s_carry[1];
    if (src.chunk_total) *src.chunk_total = s_carry[0];
    if (src.merge_parts) {
      const int mx = s_maxnch;
      *src.merge_parts = mx <= 16 ? 1 : mx <= 32 ? 2 : mx <= 64 ? 4 : 8;
    }
  }
}


}  // namespace wgbm
