// Directional Era-Splitting gain scan and split selection (gfx950).
//
// Per (era, feature) the scan walks bins in order: prefix G_L, H_L, then for
// every threshold b < B-1 the reference's gain and direction
// (warpgbm/cuda/best_split_kernel.cu:25-54).  Prefix sums and the gain are
// evaluated in float64 from the float32 histogram and rounded once to
// float32; all arithmetic here is built with -ffp-contract=off so that the
// per-node op and the batched level scan share bit-identical results.
//
// Selection (warpgbm/core.py:1322-1362): E == 1 -> crit = gain, valid iff
// crit > min_split_gain; E > 1 -> crit = (sum_e gain_e)/E in float32 (era
// order), agreement = |sum_e dir_e|, only candidates at the node's maximum
// agreement may win.  Monotone constraints require dir == required in every
// era.  Winner = lexicographic max of (agreement, crit, -flat index); a
// winner whose crit is -inf (no valid candidate) means "leaf".
#include "common.h"
#include "wgbm_internal.h"
#include "wgbm_plan.h"
#include "wgbm_cand.h"
#include <math.h>

namespace wgbm {

__device__ __forceinline__ void gain_dir(double GL, double HL, double Gt, double Ht, double mcw,
                                         double lam, float& gain, float& dir) {
  const double GR = Gt - GL;
  const double HR = Ht - HL;
  if (HL >= mcw && HR >= mcw) {
    const double a = GL * GL / (HL + lam);
    const double b = GR * GR / (HR + lam);
    const double c = Gt * Gt / (Ht + lam);
    gain = (float)((a + b) - c);
    dir = (GL / (HL + lam) > GR / (HR + lam)) ? 1.0f : -1.0f;
  } else {
    gain = 0.0f;
    dir = 0.0f;
  }
}

// Reference op: per-era gain and direction, thread per (era, feature).
__global__ void __launch_bounds__(256)
k_compute_split(const float* __restrict__ G, const float* __restrict__ H, int E, int F, int B,
                float mcw, float lam, float* __restrict__ gain, float* __restrict__ dir) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)E * F) return;
  const float* g = G + t * B;
  const float* h = H + t * B;
  double Gt = 0.0, Ht = 0.0;
  for (int b = 0; b < B; ++b) { Gt += (double)g[b]; Ht += (double)h[b]; }
  double GL = 0.0, HL = 0.0;
  float* go = gain + t * (B - 1);
  float* dO = dir + t * (B - 1);
  for (int b = 0; b < B - 1; ++b) {
    GL += (double)g[b];
    HL += (double)h[b];
    float ga, di;
    gain_dir(GL, HL, Gt, Ht, (double)mcw, (double)lam, ga, di);
    go[b] = ga;
    dO[b] = di;
  }
}

// Candidate of flat index c = f*(B-1)+b from per-era gain/dir arrays.
__device__ __forceinline__ Cand eval_cand(const float* __restrict__ gain, const float* __restrict__ dir,
                                          int E, int FB, int c, float msg, int req) {
  Cand k;
  k.idx = c;
  if (E == 1) {
    const float cr = gain[c];
    const float d = dir[c];
    const bool ok = (cr > msg) && (req == 0 || d == (float)req);
    k.agree = 0;
    k.crit = ok ? cr : -INFINITY;
  } else {
    float sg = 0.0f;
    int sd = 0;
    bool mono = true;
    for (int e = 0; e < E; ++e) {
      const float d = dir[(int64_t)e * FB + c];
      sg += gain[(int64_t)e * FB + c];
      sd += (int)d;
      mono = mono && (req == 0 || d == (float)req);
    }
    const float cr = sg / (float)E;
    k.agree = sd < 0 ? -sd : sd;
    k.crit = (cr > msg && mono) ? cr : -INFINITY;
  }
  return k;
}

// Plain-path selection for one node: out[0] = local feature, out[1] = bin,
// or (-1, -1) when no candidate is valid.
__global__ void __launch_bounds__(1024)
k_select_split(const float* __restrict__ gain, const float* __restrict__ dir, int E, int F, int B,
               float msg, const int8_t* __restrict__ mono_req, const int32_t* __restrict__ feat_idx,
               int32_t* __restrict__ out) {
  __shared__ Cand s_c[16];
  const int FB = F * (B - 1);
  Cand best = cand_none();
  for (int c = threadIdx.x; c < FB; c += blockDim.x) {
    const int f = c / (B - 1);
    const int req = mono_req ? (int)mono_req[feat_idx ? feat_idx[f] : f] : 0;
    const Cand k = eval_cand(gain, dir, E, FB, c, msg, req);
    if (cand_better(k, best)) best = k;
  }
  best = wave_best(best);
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    Cand b = s_c[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      if (cand_better(s_c[w], b)) b = s_c[w];
    if (b.idx == 0x7fffffff || b.crit == -INFINITY) { out[0] = -1; out[1] = -1; }
    else { out[0] = b.idx / (B - 1); out[1] = b.idx % (B - 1); }
  }
}

// ---------------------------------------------------------------------------
// Batched level scan, one era.  A wave scans FPW features of a node (two
// when a feature's B bins fit half a wave, else one): sub-lane b of a
// feature owns threshold b (and b + LPF, ... when B - 1 > LPF).  Each lane
// sums the bins in bin order -- the left sums over bins 0..b and the totals
// over all B bins -- so every float64 sum is the one the sequential
// per-feature loop of k_compute_split forms and the float32 gains are
// bit-identical.  When all B bins fit the feature's lanes, the total is the
// left sum of sub-lane B - 1 (the same bins added in the same order).  One
// candidate per (node, workgroup of kScanFeat waves) goes to `cand`;
// pick_node reduces them (cand_better is a strict total order, so the
// reduction order is free).
constexpr int kScanFeat = 4;   // waves per workgroup
constexpr int kScanBins32 = 32;  // bin count of the vectorised path

__host__ __device__ __forceinline__ int scan_fpw(int n_bins) { return n_bins <= 32 ? 2 : 1; }

__device__ __forceinline__ double shfl_f64(double v, int src) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __shfl((int)(uint32_t)u, src), hi = __shfl((int)(uint32_t)(u >> 32), src);
  return __builtin_bit_cast(double, ((long long)(uint32_t)hi << 32) | (uint32_t)lo);
}

__global__ void __launch_bounds__(64 * kScanFeat)
k_level_scan(const LevelScanArgs a) {
  const int B = a.n_bins;
  const int FPW = scan_fpw(B);
  const int LPF = kWave / FPW;  // lanes per feature
  const int per_wg = kScanFeat * FPW;
  const int wg_per_node = (a.n_feat + per_wg - 1) / per_wg;
  const int node = blockIdx.x / wg_per_node;
  const int wg = blockIdx.x - node * wg_per_node;
  const int lane = threadIdx.x & 63;
  const int half = lane / LPF, sl = lane - half * LPF;
  const int src0 = half * LPF;  // first lane of this lane's feature
  const int f = wg * per_wg + (threadIdx.x >> 6) * FPW + half;
  Cand best = cand_none();
  const bool active = node < a.n_nodes && a.node_active[node] != 0;
  // Every lane runs the loops (the shuffles read all lanes); lanes without a
  // feature read feature 0 and offer no candidate.
  const int fc = f < a.n_feat ? f : 0;
  const int req = a.mono_req ? (int)a.mono_req[a.feat_idx ? a.feat_idx[fc] : fc] : 0;
  const int64_t off = ((int64_t)(node < a.n_nodes ? node : 0) * a.n_feat + fc) * B;
  const float* g = a.G + off;
  const float* h = a.H + off;
  if (B == kScanBins32) {
    // 32 bins (two features per wave, lane sl owns threshold sl): every lane
    // loads its feature's 32 bins with 16-byte loads (the 32 lanes of a
    // feature read the same lines) and forms the totals and its left sums
    // itself, in bin order -- the sums the loop below forms, without a
    // shuffle per bin.
    const float4* g4 = reinterpret_cast<const float4*>(g);
    const float4* h4 = reinterpret_cast<const float4*>(h);
    float gv[kScanBins32], hv[kScanBins32];
#pragma unroll
    for (int k = 0; k < kScanBins32 / 4; ++k) {
      const float4 x = g4[k], y = h4[k];
      gv[4 * k] = x.x; gv[4 * k + 1] = x.y; gv[4 * k + 2] = x.z; gv[4 * k + 3] = x.w;
      hv[4 * k] = y.x; hv[4 * k + 1] = y.y; hv[4 * k + 2] = y.z; hv[4 * k + 3] = y.w;
    }
    const int b = sl;
    // Running sums over bins 0..q; the left sums are their values at q = b
    // (the same additions in the same order as summing bins 0..b).
    double Gt = 0.0, Ht = 0.0, GL = 0.0, HL = 0.0;
#pragma unroll
    for (int q = 0; q < kScanBins32; ++q) {
      Gt += (double)gv[q];
      Ht += (double)hv[q];
      if (q == b) { GL = Gt; HL = Ht; }
    }
    float ga, di;
    gain_dir(GL, HL, Gt, Ht, (double)a.mcw, (double)a.lam, ga, di);
    if (active && f < a.n_feat && b < B - 1) {
      const Cand k{0, (ga > a.msg && (req == 0 || di == (float)req)) ? ga : -INFINITY, f * (B - 1) + b};
      if (cand_better(k, best)) best = k;
    }
  }
  const bool total_in_lane = B <= LPF;
  for (int b0 = 0; B != kScanBins32 && b0 < B - 1; b0 += LPF) {
    const int b = b0 + sl;  // this lane's threshold
    double Gt = 0.0, Ht = 0.0, GL = 0.0, HL = 0.0;
    for (int i0 = 0; i0 < B; i0 += LPF) {
      const int iq = min(i0 + sl, B - 1);
      const float gv = g[iq], hv = h[iq];
      const int n = min(LPF, B - i0);
      for (int q = 0; q < n; ++q) {
        const double gi = (double)__shfl(gv, src0 + q), hi = (double)__shfl(hv, src0 + q);
        if (!total_in_lane) { Gt += gi; Ht += hi; }
        if (i0 + q <= b) { GL += gi; HL += hi; }
      }
    }
    if (total_in_lane) {
      Gt = shfl_f64(GL, src0 + B - 1);
      Ht = shfl_f64(HL, src0 + B - 1);
    }
    float ga, di;
    gain_dir(GL, HL, Gt, Ht, (double)a.mcw, (double)a.lam, ga, di);
    if (active && f < a.n_feat && b < B - 1) {
      const Cand k{0, (ga > a.msg && (req == 0 || di == (float)req)) ? ga : -INFINITY, f * (B - 1) + b};
      if (cand_better(k, best)) best = k;
    }
  }
  __shared__ Cand s_c[kScanFeat];
  best = wave_best(best);
  if (lane == 0) s_c[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0 && node < a.n_nodes) {
    Cand bb = s_c[0];
    for (int w = 1; w < kScanFeat; ++w)
      if (cand_better(s_c[w], bb)) bb = s_c[w];
    const int64_t o = (int64_t)node * wg_per_node + wg;
    a.cand_agree[o] = bb.agree;
    a.cand_crit[o] = bb.crit;
    a.cand_idx[o] = bb.idx;
  }
}

// Winner of one node, then (optionally) its per-era gains; run by one wave
// (lane = 0..63).
__device__ void pick_node(const LevelScanArgs& a, int node, int wg_per_node, int lane) {
  const Cand b = a.node_active[node] ? node_best(a.cand_agree, a.cand_crit, a.cand_idx, node, wg_per_node, lane)
                                     : cand_none();
  int fl, bin;
  cand_split(b, a.n_bins, fl, bin);
  if (lane == 0) {
    a.dec_feat[node] = fl;
    a.dec_bin[node] = bin;
  }
  if (a.era_gain == nullptr) return;
  const int B = a.n_bins, E = a.n_eras;
  if (E == 1) {
    // One era: its gain is the winner's criterion (the same float64 sums in
    // bin order, the same formula and rounding).
    if (lane == 0) a.era_gain[node] = fl >= 0 ? b.crit : 0.0f;
    return;
  }
  // A lane takes eras lane, lane + 64, ...: EU eras' bins (U at a time) are
  // loaded together, then each era's sums run in bin order.
  constexpr int U = 8, EU = 4;
  for (int e0 = lane; e0 < E; e0 += 64 * EU) {
    double Gt[EU], Ht[EU], GL[EU], HL[EU];
#pragma unroll
    for (int k = 0; k < EU; ++k) { Gt[k] = 0.0; Ht[k] = 0.0; GL[k] = 0.0; HL[k] = 0.0; }
    if (fl >= 0) {
      for (int q0 = 0; q0 < B; q0 += U) {
        float gv[EU][U], hv[EU][U];
#pragma unroll
        for (int k = 0; k < EU; ++k) {
          const int e = min(e0 + 64 * k, E - 1);
          const int64_t off = (((int64_t)node * E + e) * a.n_feat + fl) * B;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int q = min(q0 + u, B - 1);
            gv[k][u] = a.G[off + q];
            hv[k][u] = a.H[off + q];
          }
        }
#pragma unroll
        for (int k = 0; k < EU; ++k)
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int q = q0 + u;
            if (q < B) { Gt[k] += (double)gv[k][u]; Ht[k] += (double)hv[k][u]; }
            if (q <= bin) { GL[k] += (double)gv[k][u]; HL[k] += (double)hv[k][u]; }
          }
      }
    }
#pragma unroll
    for (int k = 0; k < EU; ++k) {
      const int e = e0 + 64 * k;
      if (e >= E) break;
      float ga = 0.0f;
      if (fl >= 0) {
        float di;
        gain_dir(GL[k], HL[k], Gt[k], Ht[k], (double)a.mcw, (double)a.lam, ga, di);
      }
      a.era_gain[(int64_t)node * E + e] = ga;
    }
  }
}

__global__ void __launch_bounds__(64)
k_level_pick(const LevelScanArgs a, int wg_per_node) {
  if (blockIdx.x < a.n_nodes) pick_node(a, blockIdx.x, wg_per_node, threadIdx.x);
}

constexpr int kDecideFusedNodes = 32;  // most nodes the fused decision launch takes

// One workgroup: the winners of every node (wave w takes nodes w, w+16, ...),
// then the leaf plan and the partition plan of the level, which depend only
// on those decisions.
__global__ void __launch_bounds__(1024)
k_decide_plans(const LevelScanArgs a, int wg_per_node, const PlanJob leaf, const PlanJob part) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int node = wv; node < a.n_nodes; node += 16) pick_node(a, node, wg_per_node, lane);
  __syncthreads();
  plan_body(leaf.src, leaf.chunk_rows, leaf.seg_first, leaf.seg_pfirst, leaf.multi, leaf.n_multi);
  __syncthreads();  // the plan's shared state is reused by the next one
  plan_body(part.src, part.chunk_rows, part.seg_first, part.seg_pfirst, part.multi, part.n_multi);
}

constexpr int kSelBlock = 256;  // (feature, threshold) candidates per workgroup of the E > 1 selection
constexpr int kGainSelEras = 4;        // eras whose bins the fused E > 1 selection loads together
constexpr int kGainSelMinNodes = 128;  // nodes from which a level takes the fused selection

int64_t level_scan_groups(int n_feat, int n_eras, int n_bins) {
  if (n_feat <= 0) return 1;
  if (n_eras == 1) return ceil_div(n_feat, kScanFeat * scan_fpw(n_bins));
  return ceil_div((int64_t)n_feat * (n_bins - 1), kSelBlock);
}

int64_t level_scan_gain_nodes(int64_t n_nodes, int n_eras, int n_feat, int n_bins, size_t cap) {
  if (n_eras == 1) return 0;
  const int64_t per = (int64_t)n_eras * n_feat * (n_bins - 1) * 5;  // float gain + int8 direction
  int64_t k = per > 0 ? (int64_t)(cap / per) : n_nodes;
  if (k < 1) k = 1;
  return k < n_nodes ? k : n_nodes;
}

// E > 1, pass 1: per-era gain and direction of every (node, era, feature,
// threshold), exactly as k_compute_split forms them (float64 prefix in bin
// order, one rounding).  Thread per (node of the group, era, feature).
__global__ void __launch_bounds__(256)
k_level_gains(const LevelScanArgs a, int node0, int n_group) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int E = a.n_eras, F = a.n_feat, B = a.n_bins;
  if (t >= (int64_t)n_group * E * F) return;
  const int jl = (int)(t / ((int64_t)E * F));
  if (!a.node_active[node0 + jl]) return;
  const int64_t off = ((int64_t)node0 * E * F + t) * B;  // histogram [node][era][feature][bin]
  const float* g = a.G + off;
  const float* h = a.H + off;
  // Bins are read U at a time (independent loads in flight), summed in bin
  // order.
  constexpr int U = 8;
  double Gt = 0.0, Ht = 0.0;
  for (int b0 = 0; b0 < B; b0 += U) {
    float gv[U], hv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { const int q = min(b0 + u, B - 1); gv[u] = g[q]; hv[u] = h[q]; }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b0 + u < B) { Gt += (double)gv[u]; Ht += (double)hv[u]; }
  }
  double GL = 0.0, HL = 0.0;
  float* go = a.era_g + t * (B - 1);
  int8_t* dO = a.era_d + t * (B - 1);
  for (int b0 = 0; b0 < B - 1; b0 += U) {
    float gv[U], hv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { const int q = min(b0 + u, B - 1); gv[u] = g[q]; hv[u] = h[q]; }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u;
      if (b < B - 1) {
        GL += (double)gv[u];
        HL += (double)hv[u];
        float ga, di;
        gain_dir(GL, HL, Gt, Ht, (double)a.mcw, (double)a.lam, ga, di);
        go[b] = ga;
        dO[b] = (int8_t)di;
      }
    }
  }
}

// E > 1, pass 2: thread per (node, feature, threshold) sums its per-era gains
// and directions in era order (eval_cand's order) and applies the monotone
// mask; one candidate per workgroup of kSelBlock thresholds.
__global__ void __launch_bounds__(kSelBlock)
k_level_sel(const LevelScanArgs a, int node0, int n_group) {
  const int E = a.n_eras, F = a.n_feat, B = a.n_bins;
  const int FB = F * (B - 1);
  const int groups = (FB + kSelBlock - 1) / kSelBlock;  // workgroups (candidate slots) per node
  const int jl = blockIdx.x / groups;
  const int wg = blockIdx.x - jl * groups;
  const int node = node0 + jl;
  const int c = wg * kSelBlock + threadIdx.x;
  Cand best = cand_none();
  if (jl < n_group && a.node_active[node] && c < FB) {
    const int f = c / (B - 1);
    const int req = a.mono_req ? (int)a.mono_req[a.feat_idx ? a.feat_idx[f] : f] : 0;
    const float* gp = a.era_g + (int64_t)jl * E * FB + c;
    const int8_t* dp = a.era_d + (int64_t)jl * E * FB + c;
    float sg = 0.0f;
    int sd = 0;
    bool mono = true;
    // U eras' loads in flight, then their sums in era order.
    constexpr int U = 8;
    for (int e0 = 0; e0 < E; e0 += U) {
      float gv[U];
      int dv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = min(e0 + u, E - 1);
        gv[u] = gp[e * FB];
        dv[u] = dp[e * FB];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e0 + u < E) {
          sg += gv[u];
          sd += dv[u];
          mono = mono && (req == 0 || dv[u] == req);
        }
      }
    }
    const float cr = sg / (float)E;
    best = Cand{sd < 0 ? -sd : sd, (cr > a.msg && mono) ? cr : -INFINITY, c};
  }
  __shared__ Cand s_c[kSelBlock / 64];
  best = wave_best(best);
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0 && jl < n_group) {
    Cand b = s_c[0];
    for (int w = 1; w < kSelBlock / 64; ++w)
      if (cand_better(s_c[w], b)) b = s_c[w];
    const int64_t o = (int64_t)node * groups + wg;
    a.cand_agree[o] = b.agree;
    a.cand_crit[o] = b.crit;
    a.cand_idx[o] = b.idx;
  }
}

// E > 1 with few bins, fused, for levels of many nodes: a thread per (node,
// feature, threshold), the candidate layout of k_level_sel.  Each era's
// bins are loaded (kGainSelEras eras at a time) and the threshold's gain and
// direction formed from the same float64 sums in bin order as k_level_gains;
// the gains are summed in era order as k_level_sel sums them.  The
// candidates are identical and the per-era gain table is never written.
template <int BM>
__global__ void __launch_bounds__(kSelBlock)
k_level_gainsel(const LevelScanArgs a, int groups) {
  const int E = a.n_eras, F = a.n_feat, B = a.n_bins;
  const int FB = F * (B - 1);
  const int node = blockIdx.x / groups;
  const int wg = blockIdx.x - node * groups;
  const int c = wg * kSelBlock + threadIdx.x;
  Cand best = cand_none();
  if (a.node_active[node] && c < FB) {
    const int f = c / (B - 1), b = c - f * (B - 1);
    const int req = a.mono_req ? (int)a.mono_req[a.feat_idx ? a.feat_idx[f] : f] : 0;
    float sg = 0.0f;
    int sd = 0;
    bool mono = true;
    const int64_t era_stride = (int64_t)F * B;
    const float* g = a.G + ((int64_t)node * E * F + f) * B;  // histogram [node][era][feature][bin]
    const float* h = a.H + ((int64_t)node * E * F + f) * B;
    constexpr int U = kGainSelEras;
    for (int e0 = 0; e0 < E; e0 += U) {
      float gv[U][BM], hv[U][BM];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t o = (int64_t)min(e0 + u, E - 1) * era_stride;
#pragma unroll
        for (int q = 0; q < BM; ++q) {
          if (q < B) {  // uniform
            gv[u][q] = g[o + q];
            hv[u][q] = h[o + q];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e0 + u >= E) break;
        double Gt = 0.0, Ht = 0.0, GL = 0.0, HL = 0.0;
#pragma unroll
        for (int q = 0; q < BM; ++q) {
          if (q < B) { Gt += (double)gv[u][q]; Ht += (double)hv[u][q]; }
          if (q <= b) { GL += (double)gv[u][q]; HL += (double)hv[u][q]; }
        }
        float ga, di;
        gain_dir(GL, HL, Gt, Ht, (double)a.mcw, (double)a.lam, ga, di);
        const int d = (int)(int8_t)di;
        sg += ga;
        sd += d;
        mono = mono && (req == 0 || d == req);
      }
    }
    const float cr = sg / (float)E;
    best = Cand{sd < 0 ? -sd : sd, (cr > a.msg && mono) ? cr : -INFINITY, c};
  }
  __shared__ Cand s_c[kSelBlock / 64];
  best = wave_best(best);
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    Cand bb = s_c[0];
    for (int w = 1; w < kSelBlock / 64; ++w)
      if (cand_better(s_c[w], bb)) bb = s_c[w];
    const int64_t o = (int64_t)node * groups + wg;
    a.cand_agree[o] = bb.agree;
    a.cand_crit[o] = bb.crit;
    a.cand_idx[o] = bb.idx;
  }
}

int level_scan_launch(const LevelScanArgs& a, hipStream_t st, const PlanJob* leaf, const PlanJob* part, bool pick) {
  if (a.n_nodes == 0) return kOk;
  if (a.n_eras == 1) {
    const int wg_per_node = (int)level_scan_groups(a.n_feat, 1, a.n_bins);
    hipLaunchKernelGGL(k_level_scan, dim3((unsigned)(a.n_nodes * (int64_t)wg_per_node)), dim3(64 * kScanFeat), 0,
                       st, a);
    WGBM_LAUNCHED("k_level_scan");
    if (!pick) return kOk;  // the level tail reduces the candidates (level_tail_fused_launch)
  } else if (a.n_bins >= 2 && a.n_bins <= 8 && a.n_nodes >= kGainSelMinNodes) {
    // many nodes: the fused selection; fewer: per-era gains, then selection
    // (200 eras x 400 features: 179-205 us against ~290 at 128 nodes, about
    // even at 64)
    const int groups = (int)level_scan_groups(a.n_feat, a.n_eras, a.n_bins);
    hipLaunchKernelGGL(k_level_gainsel<8>, dim3((unsigned)((int64_t)a.n_nodes * groups)), dim3(kSelBlock), 0, st, a,
                       groups);
    WGBM_LAUNCHED("k_level_gainsel");
  } else {
    WGBM_REQUIRE(a.era_g != nullptr && a.era_d != nullptr && a.gain_nodes >= 1, "E > 1 scan needs gain buffers");
    const int64_t groups = level_scan_groups(a.n_feat, a.n_eras, a.n_bins);
    for (int64_t n0 = 0; n0 < a.n_nodes; n0 += a.gain_nodes) {
      const int ng = (int)(a.n_nodes - n0 < a.gain_nodes ? a.n_nodes - n0 : a.gain_nodes);
      const int64_t th = (int64_t)ng * a.n_eras * a.n_feat;
      hipLaunchKernelGGL(k_level_gains, dim3((unsigned)ceil_div(th, 256)), dim3(256), 0, st, a, (int)n0, ng);
      WGBM_LAUNCHED("k_level_gains");
      hipLaunchKernelGGL(k_level_sel, dim3((unsigned)(ng * groups)), dim3(kSelBlock), 0, st, a, (int)n0, ng);
      WGBM_LAUNCHED("k_level_sel");
    }
  }
  const int groups = (int)level_scan_groups(a.n_feat, a.n_eras, a.n_bins);
  // One workgroup picks every node in the fused launch (a wave per node);
  // with more nodes, or the per-era gains of several eras to recompute, a
  // workgroup per node and the plans in launches of their own are faster
  // (200 eras, 64 nodes: 43 us against 208).
  if (leaf && part && a.n_eras == 1 && a.n_nodes <= kDecideFusedNodes) {
    hipLaunchKernelGGL(k_decide_plans, dim3(1), dim3(kPlanThreads), 0, st, a, groups, *leaf, *part);
    WGBM_LAUNCHED("k_decide_plans");
    return kOk;
  }
  hipLaunchKernelGGL(k_level_pick, dim3((unsigned)a.n_nodes), dim3(64), 0, st, a, groups);
  WGBM_LAUNCHED("k_level_pick");
  if (leaf) {
    plan_job_launch(*leaf, st);
    WGBM_LAUNCHED("k_hist_plan(leaf)");
  }
  if (part) {
    plan_job_launch(*part, st);
    WGBM_LAUNCHED("k_hist_plan(partition)");
  }
  return kOk;
}

}  // namespace wgbm

using namespace wgbm;

extern "C" int wgbm_compute_split(const float* grad_hist, const float* hess_hist, int32_t n_eras,
                                  int32_t n_feat, int32_t n_bins, float min_split_gain,
                                  float min_child_weight, float lambda, float* per_era_gain,
                                  float* per_era_direction, hipStream_t stream) {
  (void)min_split_gain;  // the reference kernel receives it but does not use it
  WGBM_REQUIRE(n_eras >= 1 && n_feat >= 0 && n_bins >= 2, "bad histogram shape [%d, %d, %d]", n_eras,
               n_feat, n_bins);
  const int64_t n = (int64_t)n_eras * n_feat;
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_compute_split, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, grad_hist,
                     hess_hist, n_eras, n_feat, n_bins, min_child_weight, lambda, per_era_gain,
                     per_era_direction);
  WGBM_LAUNCHED("k_compute_split");
  return kOk;
}

extern "C" int wgbm_select_split(const float* per_era_gain, const float* per_era_direction, int32_t n_eras,
                                 int32_t n_feat, int32_t n_bins, float min_split_gain,
                                 const int8_t* mono_required, const int32_t* feature_indices,
                                 int32_t* out_feat_bin, hipStream_t stream) {
  WGBM_REQUIRE(n_eras >= 1 && n_feat >= 0 && n_bins >= 2, "bad gain shape [%d, %d, %d]", n_eras, n_feat,
               n_bins);
  WGBM_REQUIRE((int64_t)n_feat * (n_bins - 1) < 0x7fffffff, "too many split candidates");
  hipLaunchKernelGGL(k_select_split, dim3(1), dim3(1024), 0, stream, per_era_gain, per_era_direction,
                     n_eras, n_feat, n_bins, min_split_gain, mono_required, feature_indices, out_feat_bin);
  WGBM_LAUNCHED("k_select_split");
  return kOk;
}
