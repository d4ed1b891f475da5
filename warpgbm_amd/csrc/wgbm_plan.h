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
  if (t < 3) s_carry[t] = 0;
  if (t == 0) s_maxnch = 0;
  __syncthreads();
  for (int base = 0; base < nseg; base += nt) {
    const int s = base + t;
    int v[3] = {0, 0, 0};  // chunks, is-multi, partial chunks
    if (s < nseg) {
      const int c = plan_count(src, s);
      v[0] = c < 0 ? 0 : (c == 0 ? 1 : (c + chunk_rows - 1) / chunk_rows);
      v[1] = v[0] > 1;
      v[2] = v[1] ? v[0] : 0;
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
    __syncthreads();
    int pre[3] = {0, 0, 0};
    for (int k = 0; k < wv; ++k) { pre[0] += s_sum[0][k]; pre[1] += s_sum[1][k]; pre[2] += s_sum[2][k]; }
    const int c0 = s_carry[0], c1 = s_carry[1], c2 = s_carry[2];
    if (s < nseg) {
      const int first = c0 + pre[0] + x[0] - v[0];
      const int pfirst = c2 + pre[2] + x[2] - v[2];
      seg_first[s] = first;
      if (seg_pfirst) seg_pfirst[s] = pfirst;
      if (v[1] && multi) multi[c1 + pre[1] + x[1] - 1] = s;
      if (v[1] && src.multi_desc) {
        src.multi_desc[c1 + pre[1] + x[1] - 1] = make_int4(pfirst, v[0], src.seg_slot[s], 0);
        atomicMax(&s_maxnch, v[0]);
      }
    }
    if (src.chunk_desc || src.chunk_seg) {
      // Chunk descriptors of this block of segments, all threads at once:
      // the segments' tables go to LDS, then thread t takes chunks t,
      // t + blockDim, ... and finds its segment by binary search (the last
      // segment whose first chunk is <= c; a segment without chunks has the
      // next one's first chunk).
      __shared__ int s_first[1024], s_start[1024], s_end[1024], s_slot[1024], s_pf[1024];
      if (s < nseg) {
        s_first[t] = c0 + pre[0] + x[0] - v[0];
        if (src.chunk_desc) {
          const int start = src.seg_start[s];
          s_start[t] = start;
          s_end[t] = start + max(plan_count(src, s), 0);
          s_slot[t] = src.seg_slot[s];
          s_pf[t] = v[1] ? c2 + pre[2] + x[2] - v[2] : -1;
        }
      } else {
        s_first[t] = 0x7fffffff;
      }
      __syncthreads();
      int blk_chunks = 0;
      for (int k = 0; k < nw; ++k) blk_chunks += s_sum[0][k];
      const int nb = min(nt, nseg - base);
      for (int c = c0 + t; c < c0 + blk_chunks; c += nt) {
        int lo = 0, hi = nb;  // count of segments with first <= c
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_first[mid] <= c) lo = mid + 1;
          else hi = mid;
        }
        const int j = lo - 1;
        if (src.chunk_seg) src.chunk_seg[c] = base + j;
        if (src.chunk_desc) {
          const int k = c - s_first[j];
          const int a0 = s_start[j] + k * chunk_rows;
          src.chunk_desc[c] = make_int4(a0, min(a0 + chunk_rows, s_end[j]), s_slot[j],
                                        s_pf[j] >= 0 ? s_pf[j] + k : -1);
        }
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
