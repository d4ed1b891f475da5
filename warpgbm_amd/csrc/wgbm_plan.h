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

constexpr int kPlanThreads = 1024;  // workgroup size of every plan (the kernels are launched with it)
constexpr int kPlanWaves = kPlanThreads / kWave;
constexpr int kPlanStage = 8192;    // shared words per pass (staged chunk counts, or segment tables)
constexpr int kPlanLoads = 8;       // staging loads a thread issues together
constexpr int kPlanOwnChunks = 32;  // blocked pass: a segment with more chunks is written by the whole workgroup
constexpr int kPlanBig = 256;       // such segments listed per pass
constexpr int kPlanStriped = 2;     // plans of at most this many segments per thread use the striped pass
static_assert(kPlanStage % kPlanThreads == 0 && 5 * kPlanThreads <= kPlanStage, "plan LDS layout");

// Scan state shared by the passes of one plan.
struct PlanShared {
  int32_t* sum;    // [3][waves] per-wave totals of the current pass
  int32_t* carry;  // [3] chunks, multi-chunk segments and their chunks before the pass
  int* maxnch;     // most chunks of a multi-chunk segment
  int32_t* buf;    // [kPlanStage]
};

// Exclusive workgroup scan of three per-thread values: x = inclusive wave
// scan on return, pre = totals of the lower waves.  Ends with the per-wave
// totals in sh.sum (one barrier).
__device__ __forceinline__ void plan_scan3(const PlanShared& sh, const int v[3], int x[3], int pre[3]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  x[0] = v[0]; x[1] = v[1]; x[2] = v[2];
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int y = __shfl_up(x[q], o);
      if (lane >= o) x[q] += y;
    }
  }
  constexpr int nw = kPlanWaves;
  if (lane == 63) { sh.sum[wv] = x[0]; sh.sum[nw + wv] = x[1]; sh.sum[2 * nw + wv] = x[2]; }
  __syncthreads();
  pre[0] = pre[1] = pre[2] = 0;
  for (int k = 0; k < wv; ++k) { pre[0] += sh.sum[k]; pre[1] += sh.sum[nw + k]; pre[2] += sh.sum[2 * nw + k]; }
}

// Adds the pass's totals to the carry (after the pass's last use of it).
__device__ __forceinline__ void plan_carry(const PlanShared& sh, int c0, int c1, int c2) {
  __syncthreads();
  constexpr int nw = kPlanWaves;
  if ((int)threadIdx.x == kPlanThreads - 1) {
    int tot[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < nw; ++k) { tot[0] += sh.sum[k]; tot[1] += sh.sum[nw + k]; tot[2] += sh.sum[2 * nw + k]; }
    sh.carry[0] = c0 + tot[0];
    sh.carry[1] = c1 + tot[1];
    sh.carry[2] = c2 + tot[2];
  }
  __syncthreads();
}

// Segment s's own tables: first chunk, first partial, multi-chunk list.
__device__ __forceinline__ void plan_write_segment(const PlanShared& sh, const PlanSrc& src, int s, int n, int first,
                                                   int pfirst, int mi, int32_t* __restrict__ seg_first,
                                                   int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi) {
  seg_first[s] = first;
  if (seg_pfirst) seg_pfirst[s] = pfirst;
  if (n > 1) {
    if (multi) multi[mi] = s;
    if (src.multi_desc) {
      src.multi_desc[mi] = make_int4(pfirst, n, src.seg_slot[s], 0);
      atomicMax(sh.maxnch, n);
    }
  }
}

// Striped pass over segments [base, base + kPlanThreads): thread t takes segment
// base + t.  The chunk tables are written by all threads at once: the pass's
// segment tables go to LDS, thread t takes chunks t, t + kPlanThreads, ... and
// finds its segment by binary search (the last segment whose first chunk is
// <= c; a segment without chunks has the next one's first chunk).
__device__ __forceinline__ void plan_pass_striped(const PlanShared& sh, const PlanSrc& src, int nseg, int base,
                                                  int chunk_rows, int32_t* __restrict__ seg_first,
                                                  int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi) {
  const int t = threadIdx.x;
  constexpr int nt = kPlanThreads, nw = kPlanWaves;
  const int s = base + t;
  const int n = s < nseg ? plan_chunks(src, s, chunk_rows) : 0;
  const int v[3] = {n, n > 1, n > 1 ? n : 0};
  int x[3], pre[3];
  plan_scan3(sh, v, x, pre);
  const int c0 = sh.carry[0], c1 = sh.carry[1], c2 = sh.carry[2];
  const int first = c0 + pre[0] + x[0] - v[0];
  const int pfirst = c2 + pre[2] + x[2] - v[2];
  if (s < nseg) plan_write_segment(sh, src, s, n, first, pfirst, c1 + pre[1] + x[1] - 1, seg_first, seg_pfirst, multi);
  if (src.chunk_desc || src.chunk_seg) {
    int32_t* s_first = sh.buf;
    int32_t* s_start = sh.buf + 1024;
    int32_t* s_end = sh.buf + 2048;
    int32_t* s_slot = sh.buf + 3072;
    int32_t* s_pf = sh.buf + 4096;
    if (s < nseg) {
      s_first[t] = first;
      if (src.chunk_desc) {
        const int start = src.seg_start[s];
        s_start[t] = start;
        s_end[t] = start + max(plan_count(src, s), 0);
        s_slot[t] = src.seg_slot[s];
        s_pf[t] = n > 1 ? pfirst : -1;
      }
    } else {
      s_first[t] = 0x7fffffff;
    }
    __syncthreads();
    int blk_chunks = 0;
#pragma unroll
    for (int k = 0; k < nw; ++k) blk_chunks += sh.sum[k];
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
        src.chunk_desc[c] = make_int4(a0, min(a0 + chunk_rows, s_end[j]), s_slot[j], s_pf[j] >= 0 ? s_pf[j] + k : -1);
      }
    }
  }
  plan_carry(sh, c0, c1, c2);
}

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

// Blocked pass over up to kPlanStage segments from base: their chunk counts
// are staged in LDS (coalesced loads, kPlanLoads in flight per thread), then
// thread t takes cpt consecutive ones, with one workgroup scan.  A segment of
// at most kPlanOwnChunks chunks writes its chunk tables itself; larger ones
// are listed and written by the whole workgroup.  Returns the segments done.
__device__ __forceinline__ int plan_pass_blocked(const PlanShared& sh, const PlanSrc& src, int nseg, int base,
                                                 int chunk_rows, int32_t* __restrict__ seg_first,
                                                 int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi) {
  __shared__ int s_nbig;
  __shared__ int s_big[kPlanBig][3];  // {segment, first chunk, first partial or -1}
  const int t = threadIdx.x;
  constexpr int nt = kPlanThreads;
  int32_t* s_n = sh.buf;
  const int span = min(kPlanStage, nseg - base);
  for (int k0 = 0; k0 < span; k0 += kPlanLoads * nt) {
    int nv[kPlanLoads];
#pragma unroll
    for (int u = 0; u < kPlanLoads; ++u) {
      const int k = k0 + u * nt + t;
      nv[u] = k < span ? plan_chunks(src, base + k, chunk_rows) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPlanLoads; ++u) {
      const int k = k0 + u * nt + t;
      if (k < span) s_n[k] = nv[u];
    }
  }
  if (t == 0) s_nbig = 0;
  __syncthreads();
  const int cpt = (span + nt - 1) / nt;
  const int s0 = base + t * cpt;
  const int s1 = min(s0 + cpt, base + span);
  int v[3] = {0, 0, 0};  // chunks, multi-chunk segments, their chunks
  for (int s = s0; s < s1; ++s) {
    const int n = s_n[s - base];
    v[0] += n;
    v[1] += n > 1;
    v[2] += n > 1 ? n : 0;
  }
  int x[3], pre[3];
  plan_scan3(sh, v, x, pre);
  const int c0 = sh.carry[0], c1 = sh.carry[1], c2 = sh.carry[2];
  int r0 = c0 + pre[0] + x[0] - v[0];
  int r1 = c1 + pre[1] + x[1] - v[1];
  int r2 = c2 + pre[2] + x[2] - v[2];
  const bool chunk_tables = src.chunk_desc || src.chunk_seg;
  for (int s = s0; s < s1; ++s) {
    const int n = s_n[s - base];
    const int pf = n > 1 ? r2 : -1;
    plan_write_segment(sh, src, s, n, r0, r2, r1, seg_first, seg_pfirst, multi);
    if (chunk_tables && n > 0) {
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
      plan_write_chunks(src, s, s_big[q][1], s_n[s - base], s_big[q][2], t, nt, chunk_rows);
    }
  }
  plan_carry(sh, c0, c1, c2);
  return span;
}

// One workgroup of kPlanThreads threads computes a plan.
__device__ __forceinline__ void plan_body(const PlanSrc& src, int32_t chunk_rows, int32_t* __restrict__ seg_first,
                          int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi,
                          int32_t* __restrict__ n_multi) {
  __shared__ int32_t s_sum[3 * kPlanWaves];
  __shared__ int32_t s_carry[3];
  __shared__ int s_maxnch;
  __shared__ int32_t s_buf[kPlanStage];
  const PlanShared sh{s_sum, s_carry, &s_maxnch, s_buf};
  const int nseg = src.nseg_dev ? *src.nseg_dev : src.nseg;
  if (src.nseg_out && threadIdx.x == 0) *src.nseg_out = nseg;
  const int t = threadIdx.x, lane = t & 63;
  constexpr int nt = kPlanThreads;
  if (src.chunk_out) {
    // Rows per chunk for this launch: enough chunks to give every slot a
    // work item (chunk_target of them), within [chunk_min, chunk_rows].  With
    // C >= rows / tgt, the live segments' chunks number sum ceil(count / C)
    // < rows / C + live <= tgt + live, so tgt = chunk_target - live + 1 (at
    // least half the target) keeps the launch within chunk_target chunks:
    // no slot takes a second work item after the others have finished.
    __shared__ int s_rows, s_live;
    if (t == 0) { s_rows = 0; s_live = 0; }
    __syncthreads();
    int mine = 0, live = 0;
    for (int s = t; s < nseg; s += nt) {
      const int c = max(plan_count(src, s), 0);
      mine += c;
      live += c > 0;
    }
    for (int o = 32; o >= 1; o >>= 1) { mine += __shfl_xor(mine, o); live += __shfl_xor(live, o); }
    if (lane == 0) { atomicAdd(&s_rows, mine); atomicAdd(&s_live, live); }
    __syncthreads();
    const int tgt = max(src.chunk_target - s_live + 1, (src.chunk_target + 1) / 2);
    const int want = (s_rows + tgt - 1) / tgt;
    chunk_rows = min(max((want + 63) & ~63, src.chunk_min), chunk_rows);  // whole 64-row blocks
    if (t == 0) *src.chunk_out = chunk_rows;
  }
  if (t < 3) s_carry[t] = 0;
  if (t == 0) s_maxnch = 0;
  __syncthreads();
  // Few segments per thread: striped passes (one dependent load round, the
  // chunk tables from LDS); many: blocked passes over LDS-staged counts.
  if (nseg <= kPlanStriped * nt) {
    for (int base = 0; base < nseg; base += nt)
      plan_pass_striped(sh, src, nseg, base, chunk_rows, seg_first, seg_pfirst, multi);
  } else {
    for (int base = 0; base < nseg;)
      base += plan_pass_blocked(sh, src, nseg, base, chunk_rows, seg_first, seg_pfirst, multi);
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
