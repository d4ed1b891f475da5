// Batched level-wise boosting on one stream (gfx950).
//
// One call trains n_rounds boosting rounds with no host synchronisation:
// per round the objective kernel writes (g, h) in root order, then for every
// tree level d = 0..D the level's kernels run for ALL frontier nodes of ALL
// K class trees at once:
//   histogram of the smaller children (one launch; root: every era of every
//   class) -> sibling = parent - built -> fused gain scan + DES selection ->
//   leaf values and score updates of the level's leaves -> partition of the
//   split nodes into the next level.
// The per-node decisions and leaf values are recorded into level-order
// (heap id) node tables, from which the host builds the reference's nested
// tree dicts once training returns.
//
// Trees are identical to the plain per-node path: both use the same
// histogram kernel (same chunking), the same gain/selection arithmetic, the
// same stable partition and the same leaf reduction.  The reference's DFS
// growth (warpgbm/core.py:1364-1470) freezes g/h per round (:1545-1547), so
// level order and DFS order produce the same tree.
#include "common.h"
#include "wgbm_internal.h"
#include "wgbm_level.h"
#include "wgbm_stats.h"
#include "wgbm_plan.h"
#include "../../include/warpgbm_amd.h"
#include <algorithm>

namespace wgbm {

struct FitLayout {
  size_t total = 0;
  // offsets
  size_t perm_root, gh_root, perm_pp[2], gh_pp[2];
  size_t hg[2], hh[2];
  struct Tables { size_t seg_start, seg_count, node_active, node_start, node_count, stats; } tab[2];
  size_t dec_feat, dec_bin, era_gain, cand_agree, cand_crit, cand_idx, era_g, era_d;
  int64_t gain_nodes;
  size_t b_seg_start, b_seg_count, b_seg_slot, n_build, sib, n_split;
  size_t h_seg_first, h_seg_pfirst, h_nseg, h_multi, h_n_multi, h_chunk, h_part, h_pflag, h_fine, h_desc, h_total, h_mdesc,
      h_mparts, loss_part;
  size_t stat_part, round_ctr, fidx_cur;
  size_t p_flag, p_seg_first, p_nseg, p_multi, p_n_multi, p_lbase, p_rbase, p_cscan, p_segL, p_segP, codes_t, p_cseg;
  size_t l_seg_first, l_nseg, l_multi, l_n_multi, l_part, l_val;
  size_t tree_stats, curs[2], child_part, row_leaf;
  size_t root_seg_start, root_seg_count, root_seg_slot, root_nseg, era_start;
  // sizes
  int64_t KN, max_nodes, node_elems, hist_nodes, part_chunks, leaf_chunks, hist_part_floats;
  HistConfig hc;
};

static bool fit_one_plane(const wgbm_fit_config& c) { return c.objective == 0 && c.const_weight != 0; }

static HistConfig fit_hist_config(int n_feat, int n_bins, int64_t n_cols, bool one_plane) {
  return pick_hist_config(n_feat, n_bins, n_cols, one_plane);
}

static FitLayout fit_layout(const wgbm_fit_config& c) {
  FitLayout L;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t r = o; o = al(o + bytes); return r; };
  const int64_t K = c.n_classes, N = c.n_rows, E = c.n_eras, D = c.max_depth;
  L.KN = K * N;
  L.max_nodes = K << D;                       // nodes of the deepest level
  L.node_elems = E * (int64_t)c.n_feat * c.n_bins;
  L.hist_nodes = D >= 1 ? (K << (D - 1)) : K;  // deepest level with histograms
  L.hc = fit_hist_config(c.n_feat, c.n_bins, c.n_cols, fit_one_plane(c));
  L.perm_root = take(4 * L.KN);
  L.gh_root = take(8 * L.KN);
  for (int i = 0; i < 2; ++i) { L.perm_pp[i] = take(4 * L.KN); L.gh_pp[i] = take(8 * L.KN); }
  for (int i = 0; i < 2; ++i) {
    L.hg[i] = take(4 * L.hist_nodes * L.node_elems);
    L.hh[i] = take(4 * L.hist_nodes * L.node_elems);
  }
  const int64_t MS = L.max_nodes * E;
  for (int i = 0; i < 2; ++i) {
    L.tab[i].seg_start = take(4 * MS);
    L.tab[i].seg_count = take(4 * MS);
    L.tab[i].node_active = take(4 * L.max_nodes);
    L.tab[i].node_start = take(4 * L.max_nodes);
    L.tab[i].node_count = take(4 * L.max_nodes);
    L.tab[i].stats = take(4 * kStatWords * L.max_nodes);
  }
  const int64_t wgpn = level_scan_groups(c.n_feat, c.n_eras, c.n_bins);
  // E > 1: per-era gains of as many nodes as fit 1 GiB, scanned in groups
  L.gain_nodes = level_scan_gain_nodes(L.hist_nodes, c.n_eras, c.n_feat, c.n_bins, (size_t)1 << 30);
  const int64_t gvals = L.gain_nodes * E * c.n_feat * (int64_t)(c.n_bins - 1);
  L.era_g = take(4 * gvals);
  L.era_d = take(gvals);
  L.dec_feat = take(4 * L.max_nodes);
  L.dec_bin = take(4 * L.max_nodes);
  L.era_gain = take(4 * MS);
  L.cand_agree = take(4 * L.max_nodes * wgpn);
  L.cand_crit = take(4 * L.max_nodes * wgpn);
  L.cand_idx = take(4 * L.max_nodes * wgpn);
  L.b_seg_start = take(4 * MS);
  L.b_seg_count = take(4 * MS);
  L.b_seg_slot = take(4 * MS);
  L.n_build = take(4);
  L.sib = take(4 * 3 * L.max_nodes);
  L.n_split = take(4);
  L.h_seg_first = take(4 * (MS + 1));
  L.h_seg_pfirst = take(4 * (MS + 1));
  L.h_nseg = take(4);
  L.h_multi = take(4 * hist_max_multi(L.hc, c.n_feat, L.KN));
  L.h_n_multi = take(4);
  L.h_chunk = take(4);
  // chunk descriptors: enough for any level (every segment ragged)
  L.h_desc = take(sizeof(int4) * hist_max_chunks(L.hc, c.n_feat, L.KN, MS));
  L.h_total = take(4);
  L.h_mdesc = take(sizeof(int4) * hist_max_multi(L.hc, c.n_feat, L.KN));
  L.h_mparts = take(4);
  L.loss_part = take(8 * 2 * kStatBlocks);
  L.hist_part_floats = (int64_t)hist_partial_words(L.hc, c.n_feat, c.n_bins, L.KN);
  L.h_part = take(sizeof(longlong2) * L.hist_part_floats);
  L.h_pflag = take(4 * hist_partial_flags(L.hc, c.n_feat, L.KN));
  L.h_fine = take(8 * L.hist_nodes * L.node_elems);  // fine sums: one slot per parent (and per root) node-era
  L.part_chunks = ceil_div(L.KN, kPartChunk) + MS;
  L.p_flag = take(L.KN);
  L.p_seg_first = take(4 * (MS + 1));
  L.p_nseg = take(4);
  L.p_multi = take(4 * (L.KN / (kPartChunk + 1) + 2));
  L.p_n_multi = take(4);
  L.p_lbase = take(4 * L.part_chunks);
  L.p_cseg = take(4 * L.part_chunks);
  L.p_rbase = take(4 * L.part_chunks);
  L.p_cscan = take(4 * (L.part_chunks + 1));
  L.codes_t = take((size_t)N * c.n_cols);  // column-major copy of the bin matrix
  L.p_segL = take(4 * MS);
  L.p_segP = take(4 * 2 * (MS + 1));
  // statistics partials: per block of the objective (K classes) or per
  // partition chunk and child side
  const int64_t stat_parts = (int64_t)K * kStatBlocks > 2 * L.part_chunks ? (int64_t)K * kStatBlocks : 2 * L.part_chunks;
  L.stat_part = take(4 * kStatWords * stat_parts);
  L.round_ctr = take(4 * 2);           // {round being built, next round}
  L.fidx_cur = take(4 * (int64_t)c.n_feat);  // this round's feature subset
  L.leaf_chunks = leaf_max_chunks(L.KN, L.max_nodes);
  L.l_seg_first = take(4 * (L.max_nodes + 1));
  L.l_nseg = take(4);
  L.l_multi = take(4 * (L.leaf_chunks + 1));
  L.l_n_multi = take(4);
  L.l_part = take(sizeof(I128) * 2 * L.leaf_chunks);
  L.tree_stats = take(4 * kStatWords * K);    // per class: its root's statistics (scale of the leaf sums)
  // split cursors of the fused levels (at most kScanLdsSeg nodes), by level parity
  for (int i = 0; i < 2; ++i) L.curs[i] = take(8 * kCurStride * (size_t)std::min<int64_t>(L.max_nodes, kScanLdsSegMax));
  L.child_part = take(sizeof(I128) * 4 * L.part_chunks);  // last split level: children's sums per chunk
  L.row_leaf = take(2 * (size_t)N);                        // each row's leaf in the round's tree (one class)
  L.l_val = take(4 * L.max_nodes);
  L.root_seg_start = take(4 * K * E);
  L.root_seg_count = take(4 * K * E);
  L.root_seg_slot = take(4 * K * E);
  L.root_nseg = take(4);
  L.era_start = take(4 * E);
  L.total = o;
  return L;
}

// Start of a round: advance the round counter, copy the round's feature
// subset, and write the root tables (class c occupies positions
// [c*N, (c+1)*N) in root order).  The round index lives on the device so
// that one captured round can be replayed for every round.
struct RootArgs {
  const int32_t* era_count;
  int E, K, N;
  int32_t *era_start, *seg_start, *seg_count, *seg_slot, *nseg;
  int32_t *t_seg_start, *t_seg_count, *t_active, *t_node_start, *t_node_count;
  const int32_t* round_ctr;
  const int32_t* feat_all;
  int n_feat;
  int32_t* fidx_cur;
};

// One workgroup: the root segment tables, level-0 node tables and feature
// subset of round round_ctr[1] (the round about to be built; the counter
// itself advances later in the round, in k_root_stats_plan).
__device__ __forceinline__ void root_tables_block(const RootArgs& a) {
  const int E = a.E, K = a.K, N = a.N;
  const int r = a.round_ctr[1];
  if (a.feat_all)
    for (int i = threadIdx.x; i < a.n_feat; i += blockDim.x) a.fidx_cur[i] = a.feat_all[(int64_t)r * a.n_feat + i];
  int run = 0;  // era starts: exclusive scan of the era counts
  for (int base = 0; base < E; base += blockDim.x) {
    const int e = base + threadIdx.x;
    int tot;
    const int ex = block_exscan(e < E ? a.era_count[e] : 0, &tot);  // syncs: every thread has read round_ctr[1]
    if (e < E) a.era_start[e] = run + ex;
    run += tot;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < K * E; s += blockDim.x) {
    const int c = s / E, e = s - c * E;
    const int st = c * N + a.era_start[e];
    a.seg_start[s] = st;
    a.seg_count[s] = a.era_count[e];
    a.seg_slot[s] = s;
    a.t_seg_start[s] = st;
    a.t_seg_count[s] = a.era_count[e];
  }
  for (int c = threadIdx.x; c < K; c += blockDim.x) {
    a.t_active[c] = 1;
    a.t_node_start[c] = c * N;
    a.t_node_count[c] = run;
  }
  if (threadIdx.x == 0) *a.nseg = K * E;
}

// codes_t[c][r] = codes[r][c], through 64 x 64 LDS tiles (row-contiguous
// reads, column-contiguous writes).
__global__ void __launch_bounds__(256)
k_transpose_codes(const int8_t* __restrict__ codes, int64_t n_rows, int64_t n_cols, int8_t* __restrict__ codes_t) {
  __shared__ int8_t tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int t = threadIdx.x;
  {
    const int r = t >> 2, q = (t & 3) * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t rr = r0 + r, cc = c0 + q + k;
      tile[r][q + k] = (rr < n_rows && cc < n_cols) ? codes[rr * n_cols + cc] : (int8_t)0;
    }
  }
  __syncthreads();
  {
    const int c = t >> 2, q = (t & 3) * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t rr = r0 + q + k, cc = c0 + c;
      if (rr < n_rows && cc < n_cols) codes_t[cc * n_rows + rr] = tile[q + k][c];
    }
  }
}

// Same transpose with 4-byte loads and stores (n_cols % 4 == 0, n_rows % 4
// == 0, 4-byte aligned): tile rows are read as words, and each output word
// packs one column's codes of 4 consecutive rows.
__global__ void __launch_bounds__(256)
k_transpose_codes4(const int8_t* __restrict__ codes, int64_t n_rows, int64_t n_cols, int8_t* __restrict__ codes_t) {
  // [row][word]: 64 rows x 64 codes; word w of row r sits in column
  // w ^ ((r >> 2) & 15), so the 16 row groups a wave reads in the second
  // phase (rows 4q + i, one word column) fall in 16 different banks.
  __shared__ uint32_t tile[64][16];
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int t = threadIdx.x;
  const int w = t & 15;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int r = (t >> 4) + 16 * k;
    const int64_t rr = r0 + r, cc = c0 + 4 * w;
    tile[r][w ^ ((r >> 2) & 15)] =
        (rr < n_rows && cc < n_cols) ? *reinterpret_cast<const uint32_t*>(codes + rr * n_cols + cc) : 0u;
  }
  __syncthreads();
  const int q = t & 15;  // output word: rows 4q .. 4q+3
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int c = (t >> 4) + 16 * k;  // column within the tile
    const int64_t cc = c0 + c, rr = r0 + 4 * q;
    if (cc < n_cols && rr < n_rows) {
      const int sh = 8 * (c & 3);
      const int wc = (c >> 2) ^ q;  // rows 4q .. 4q+3 share the swizzle q
      const uint32_t b0 = (tile[4 * q + 0][wc] >> sh) & 0xffu;
      const uint32_t b1 = (tile[4 * q + 1][wc] >> sh) & 0xffu;
      const uint32_t b2 = (tile[4 * q + 2][wc] >> sh) & 0xffu;
      const uint32_t b3 = (tile[4 * q + 3][wc] >> sh) & 0xffu;
      *reinterpret_cast<uint32_t*>(codes_t + cc * n_rows + rr) = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
    }
  }
}

// Same transpose over 128 x 128 tiles with 8-byte loads and 16-byte stores
// (n_cols % 8 == 0, n_rows % 16 == 0, codes 8-byte and codes_t 16-byte
// aligned): each thread has 64 bytes in flight each way, and eight
// consecutive threads write one column's 128 rows, a whole line.  Thread
// (w, g) gathers word column w of rows 16g .. 16g + 15 and writes the four
// columns 4w .. 4w + 3 of those rows.  Tiles are numbered row tile major
// and dealt to the XCDs in contiguous runs (workgroup b runs on XCD b % 8),
// so the column tiles of a row tile -- whose rows share lines when a row
// is not a whole number of lines -- are read one after another through
// one L2.
constexpr int kTrTile = 128;
constexpr int kTrRowWords = kTrTile / 4 + 1;  // a tile row in LDS, padded by one word
__global__ void __launch_bounds__(256)
k_transpose_codes16(const int8_t* __restrict__ codes, int64_t n_rows, int64_t n_cols, int8_t* __restrict__ codes_t,
                    int64_t n_ctiles, int64_t per_xcd) {
  __shared__ uint32_t tile[kTrTile * kTrRowWords];
  const int64_t tid = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int64_t rt = tid / n_ctiles;
  if (rt * kTrTile >= n_rows) return;  // past the last tile (uniform)
  const int64_t r0 = rt * kTrTile, c0 = (tid - rt * n_ctiles) * kTrTile;
  const int t = threadIdx.x;
  const int ch = t & 15;  // 8-byte chunk of a tile row
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int r = (t >> 4) + 16 * k;
    const int64_t rr = r0 + r, cc = c0 + 8 * ch;
    uint2 v = make_uint2(0u, 0u);
    if (rr < n_rows && cc < n_cols) v = *reinterpret_cast<const uint2*>(codes + rr * n_cols + cc);
    tile[r * kTrRowWords + 2 * ch] = v.x;
    tile[r * kTrRowWords + 2 * ch + 1] = v.y;
  }
  __syncthreads();
  const int w = t >> 3, g = t & 7;
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = tile[(16 * g + i) * kTrRowWords + w];
  const int64_t rr = r0 + 16 * g;
  if (rr >= n_rows) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t cc = c0 + 4 * w + j;
    if (cc >= n_cols) break;
    const int sh = 8 * j;
    uint32_t o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      o[q] = ((x[4 * q] >> sh) & 0xffu) | (((x[4 * q + 1] >> sh) & 0xffu) << 8) |
             (((x[4 * q + 2] >> sh) & 0xffu) << 16) | (((x[4 * q + 3] >> sh) & 0xffu) << 24);
    *reinterpret_cast<uint4*>(codes_t + cc * n_rows + rr) = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

__global__ void __launch_bounds__(256)
k_root_perm(const int32_t* __restrict__ order, int64_t N, int K, int32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * K) return;
  const int64_t p = i % N;
  perm[i] = order ? order[p] : (int32_t)p;
}

// A recorded tree walked by rows: class c's tree of round round_ctr[0] in
// the node tables (heap order, 2^(D+1) - 1 nodes).  A workgroup stages the
// tree's nodes in LDS when they fit; rows go left when code <= bin, as the
// partition sends them, and the walk stops at a leaf.
struct TreeArgs {
  const int32_t* status;
  const int32_t* feat;
  const int32_t* bin;
  const float* value;
  const int32_t* round_ctr;
  int K, D;
  const int8_t* codes_t;  // column-major codes [columns][N]
  int64_t N;
  float lr;
  // One class only: each row's leaf (heap index) in the tree, written by
  // the tree's partition; null to walk the tree from the codes.
  const int16_t* row_leaf;
};

constexpr int kApplyRows = 4;      // rows a thread walks together (their code loads in flight at once)
constexpr int kApplyStage = 2047;  // nodes staged (trees of depth <= 10)

struct TreeStage {
  int* f;    // per node: split column, -1 leaf, -2 absent
  int* b;
  float* v;
  int64_t o;  // the tree's first node in the tables
  bool staged;
};

// Every thread of the workgroup calls it (one barrier when staged).
__device__ __forceinline__ TreeStage tree_stage(const TreeArgs& t, int c, int* sf, int* sb, float* sv) {
  const int NH = (2 << t.D) - 1;
  TreeStage ts{sf, sb, sv, ((int64_t)t.round_ctr[0] * t.K + c) * NH, NH <= kApplyStage};
  if (ts.staged) {
    for (int h = threadIdx.x; h < NH; h += blockDim.x) {
      const int st = t.status[ts.o + h];
      sf[h] = st == 1 ? t.feat[ts.o + h] : (st == 2 ? -1 : -2);
      sb[h] = t.bin[ts.o + h];
      sv[h] = t.value[ts.o + h];
    }
    __syncthreads();
  }
  return ts;
}

__device__ __forceinline__ int tree_feat(const TreeArgs& t, const TreeStage& ts, int h) {
  if (ts.staged) return ts.f[h];
  const int st = t.status[ts.o + h];
  return st == 1 ? t.feat[ts.o + h] : (st == 2 ? -1 : -2);
}

// The leaf value (the raw leaf, before the learning rate) of each of U
// rows, or 0 with ok false for rows past N or that reach no leaf.
template <int U>
__device__ __forceinline__ void tree_walk(const TreeArgs& t, const TreeStage& ts, const int64_t (&row)[U],
                                          float (&val)[U], bool (&ok)[U]) {
  if (t.row_leaf) {  // the leaves the partition recorded
    const int NH = (2 << t.D) - 1;
    int h[U];
#pragma unroll
    for (int u = 0; u < U; ++u) h[u] = row[u] < t.N ? (int)t.row_leaf[row[u]] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = h[u] >= 0 && h[u] < NH && tree_feat(t, ts, h[u]) == -1;
      val[u] = ok[u] ? (ts.staged ? ts.v[h[u]] : t.value[ts.o + h[u]]) : 0.0f;
    }
    return;
  }
  int h[U];
  bool done[U];
#pragma unroll
  for (int u = 0; u < U; ++u) { h[u] = 0; done[u] = row[u] >= t.N; }
  for (int d = 0; d < t.D; ++d) {
    int8_t code[U];
    int f[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      f[u] = done[u] ? -1 : tree_feat(t, ts, h[u]);
      code[u] = f[u] >= 0 ? t.codes_t[(int64_t)f[u] * t.N + row[u]] : (int8_t)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (f[u] >= 0) {
        const int b = ts.staged ? ts.b[h[u]] : t.bin[ts.o + h[u]];
        h[u] = 2 * h[u] + 1 + ((int)code[u] <= b ? 0 : 1);
      } else {
        done[u] = true;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    ok[u] = row[u] < t.N && tree_feat(t, ts, h[u]) == -1;
    val[u] = ok[u] ? (ts.staged ? ts.v[h[u]] : t.value[ts.o + h[u]]) : 0.0f;
  }
}

// Regression gradients in root position order, with one statistics partial
// per block (workgroups [0, nblk), grid-stride over positions, kApplyRows at
// a time), and the root tables (workgroup nblk).  With apply, the previous
// round's tree is first added to each row's score (f32(lr) * leaf, the
// update k_apply_tree makes) -- the walk and the gradient of a row in one
// pass.
__global__ void __launch_bounds__(256)
k_round_start_regression(const float* __restrict__ y, float* __restrict__ pred, const float* __restrict__ w,
                         int64_t n, const int32_t* __restrict__ perm0, float2* __restrict__ gh,
                         int32_t* __restrict__ part, int nblk, const TreeArgs ta, int apply, const RootArgs ra) {
  if ((int)blockIdx.x == nblk) {  // the last workgroup writes the round's root tables
    root_tables_block(ra);
    return;
  }
  __shared__ int s_f[kApplyStage], s_b[kApplyStage];
  __shared__ float s_v[kApplyStage];
  const int blk = blockIdx.x;
  StatAcc acc;
  const int64_t T = (int64_t)nblk * blockDim.x;
  int64_t p0 = (int64_t)blk * blockDim.x + threadIdx.x;
  int64_t row[kApplyRows];
  auto load_rows = [&]() {
#pragma unroll
    for (int u = 0; u < kApplyRows; ++u) {
      const int64_t p = p0 + u * T;
      row[u] = p < n ? (int64_t)perm0[p] : n;  // n: no row
    }
  };
  load_rows();  // the first rows' ids are in flight while the tree is staged
  TreeStage ts{};
  if (apply) ts = tree_stage(ta, 0, s_f, s_b, s_v);
  for (; p0 < n; p0 += T * kApplyRows, load_rows()) {
    // the rows' score, target and weight, loaded before the walk so that
    // they are in flight with its loads (a position past n reads row 0)
    float f[kApplyRows], yv[kApplyRows], wv[kApplyRows];
#pragma unroll
    for (int u = 0; u < kApplyRows; ++u) {
      const int64_t r = row[u] < n ? row[u] : 0;
      f[u] = pred[r];
      yv[u] = y[r];
      wv[u] = w[r];
    }
    float lv[kApplyRows];
    bool ok[kApplyRows];
    if (apply) {
      tree_walk<kApplyRows>(ta, ts, row, lv, ok);
    } else {
#pragma unroll
      for (int u = 0; u < kApplyRows; ++u) { lv[u] = 0.0f; ok[u] = false; }
    }
#pragma unroll
    for (int u = 0; u < kApplyRows; ++u) {
      const int64_t p = p0 + u * T;
      if (p >= n) continue;
      if (ok[u]) {
        f[u] = f[u] + ta.lr * lv[u];
        pred[row[u]] = f[u];
      }
      const float g = (yv[u] - f[u]) * wv[u];
      gh[p] = make_float2(g, wv[u]);
      acc.add(g, wv[u]);
    }
  }
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(part + (int64_t)blk * kStatWords);
}

// Round start, multiclass: softmax gradients (workgroups [0, nblk)) and the
// root tables (workgroup nblk).
__global__ void __launch_bounds__(256)
k_round_start_softmax(const float* __restrict__ F, const int32_t* __restrict__ y, const float* __restrict__ w,
                      int64_t n, int K, const int32_t* __restrict__ perm0, float2* __restrict__ gh, int nblk,
                      const RootArgs ra) {
  if ((int)blockIdx.x == nblk) {
    root_tables_block(ra);
    return;
  }
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = p0 < n;
  const int64_t p = ok ? p0 : n - 1;
  const int64_t r = perm0[p];
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
    if (ok) gh[(int64_t)k * n + p] = make_float2(gk, hk);
  }
}


// Training loss after a round: per-block float64 partials of (sum w * loss,
// sum w), then one workgroup folds them and writes train_loss[round].
// Regression: loss = (y - s)^2.  Multiclass: loss = -log(clamp(softmax(s)_y,
// 1e-15, 1 - 1e-15)), the softmax formed as the gradient kernel forms it.
__global__ void __launch_bounds__(256)
k_loss_part(const float* __restrict__ y, const int32_t* __restrict__ labels, const float* __restrict__ s,
            const float* __restrict__ w, int64_t n, int K, double* __restrict__ part) {
  double num = 0.0, den = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const double wr = (double)w[r];
    double loss;
    if (labels == nullptr) {
      const double e = (double)(y[r] - s[r]);
      loss = e * e;
    } else {
      const float* f = s + r * K;
      float mx = f[0];
      for (int k = 1; k < K; ++k) mx = fmaxf(mx, f[k]);
      float den_s = 0.0f;
      for (int k = 0; k < K; ++k) den_s += expf(f[k] - mx);
      const float p = fminf(fmaxf(expf(f[labels[r]] - mx) / den_s, 1e-15f), 1.0f - 1e-15f);
      loss = -(double)logf(p);
    }
    num += wr * loss;
    den += wr;
  }
  __shared__ double sn[256], sd[256];
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) { sn[threadIdx.x] += sn[threadIdx.x + o]; sd[threadIdx.x] += sd[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = sn[0]; part[2 * blockIdx.x + 1] = sd[0]; }
}

__global__ void __launch_bounds__(256)
k_loss_final(const double* __restrict__ part, int nb, const int32_t* __restrict__ round_ctr, float* __restrict__ out) {
  __shared__ double sn[256], sd[256];
  double num = 0.0, den = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) { num += part[2 * b]; den += part[2 * b + 1]; }
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) { sn[threadIdx.x] += sn[threadIdx.x + o]; sd[threadIdx.x] += sd[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[round_ctr[0]] = (float)(sn[0] / sd[0]);
}

// The root statistics of every class (workgroups [0, K)) and the root
// histogram plan (workgroup K) in one launch of 1024-thread workgroups.
// The root statistics are also the round's tree statistics (the scale of
// its leaf sums); workgroup 0 advances the round counter (the round start
// before it reads the previous round's index); workgroup K clears level 0's
// split cursors and runs the root histogram plan.
__global__ void __launch_bounds__(1024)
k_root_stats_plan(const int32_t* __restrict__ part, int n_part, int32_t* __restrict__ out,
                  int32_t* __restrict__ tree_out, int K, const PlanJob j, unsigned long long* __restrict__ cur0,
                  int n_cur, int32_t* __restrict__ round_ctr) {
  if ((int)blockIdx.x == K) {
    for (int q = threadIdx.x; q < n_cur; q += blockDim.x) cur0[(int64_t)q * kCurStride] = 0ull;
    plan_body(j.src, j.chunk_rows, j.seg_first, j.seg_pfirst, j.multi, j.n_multi);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int r = round_ctr[1];
    round_ctr[0] = r;
    round_ctr[1] = r + 1;
  }
  const int32_t* p = part + (int64_t)blockIdx.x * n_part * kStatWords;
  StatAcc acc;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) acc.load_merge(p + (int64_t)i * kStatWords);
  acc.block_reduce();
  if (threadIdx.x == 0) {
    acc.store(out + (int64_t)blockIdx.x * kStatWords);
    acc.store(tree_out + (int64_t)blockIdx.x * kStatWords);
  }
}

// The round's trees added to the scores: thread q of class c takes rows
// q, q + T, ... (T threads per class), kApplyRows walked together (or
// their recorded leaves read), and adds f32(lr) * leaf (each row's one
// update of the round, the value the per-node engine adds).
__global__ void __launch_bounds__(256)
k_apply_tree(const TreeArgs ta, float* __restrict__ scores, int blocks_per_class) {
  const int c = blockIdx.x / blocks_per_class;
  const int blk = blockIdx.x - c * blocks_per_class;
  __shared__ int s_f[kApplyStage], s_b[kApplyStage];
  __shared__ float s_v[kApplyStage];
  const TreeStage ts = tree_stage(ta, c, s_f, s_b, s_v);
  const int64_t T = (int64_t)blocks_per_class * blockDim.x;
  for (int64_t r0 = (int64_t)blk * blockDim.x + threadIdx.x; r0 < ta.N; r0 += T * kApplyRows) {
    int64_t row[kApplyRows];
#pragma unroll
    for (int u = 0; u < kApplyRows; ++u) row[u] = r0 + u * T;
    float v[kApplyRows];
    bool ok[kApplyRows];
    tree_walk<kApplyRows>(ta, ts, row, v, ok);
#pragma unroll
    for (int u = 0; u < kApplyRows; ++u) {
      if (ok[u]) {
        float* d = scores + row[u] * ta.K + c;
        *d = *d + ta.lr * v[u];
      }
    }
  }
}

}  // namespace wgbm

using namespace wgbm;

static bool valid_slots(int32_t slots) {
  return slots == 0 || slots == 128 || slots == 256 || slots == 512 || slots == 1024;
}

extern "C" int wgbm_set_hist_config(int32_t chunk_rows, int32_t waves, int32_t features_per_lane,
                                    int32_t rows_in_flight, int32_t slots) {
  WGBM_REQUIRE(chunk_rows >= 0 && (waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8 || waves == 16),
               "bad histogram configuration");
  WGBM_REQUIRE(features_per_lane == 0 || features_per_lane == 1 || features_per_lane == 2 ||
                   features_per_lane == 4 || features_per_lane == 8 || features_per_lane == 16,
               "features_per_lane must be 0 (auto), 1, 2, 4, 8 or 16");
  WGBM_REQUIRE(rows_in_flight >= 0 && rows_in_flight <= 64, "rows_in_flight must be in [0, 64]");
  WGBM_REQUIRE(valid_slots(slots), "slots must be 0 (auto), 128, 256, 512 or 1024");
  HistConfig c;
  c.chunk_rows = chunk_rows;
  c.waves = waves;
  c.fpl = features_per_lane;
  c.depth = rows_in_flight;
  c.slots = slots;
  set_hist_override(c);
  return kOk;
}

extern "C" int wgbm_set_hist_config_shape(int32_t n_feat, int32_t n_bins, int64_t n_cols, int32_t const_hessian,
                                          int32_t chunk_rows, int32_t waves, int32_t features_per_lane,
                                          int32_t rows_in_flight, int32_t slots) {
  WGBM_REQUIRE(n_feat >= 1 && n_bins >= 1 && n_bins <= 127 && n_cols >= 1, "bad histogram shape");
  WGBM_REQUIRE(chunk_rows >= 0 && (waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8 || waves == 16),
               "bad histogram configuration");
  WGBM_REQUIRE(features_per_lane == 0 || features_per_lane == 1 || features_per_lane == 2 ||
                   features_per_lane == 4 || features_per_lane == 8 || features_per_lane == 16,
               "features_per_lane must be 0 (auto), 1, 2, 4, 8 or 16");
  WGBM_REQUIRE(rows_in_flight >= 0 && rows_in_flight <= 64, "rows_in_flight must be in [0, 64]");
  WGBM_REQUIRE(valid_slots(slots), "slots must be 0 (auto), 128, 256, 512 or 1024");
  HistConfig c;
  c.chunk_rows = chunk_rows;
  c.waves = waves;
  c.fpl = features_per_lane;
  c.depth = rows_in_flight;
  c.slots = slots;
  if (c.fpl > 0) {
    WGBM_REQUIRE(lds_bytes(c, n_bins, const_hessian != 0) <= 159 * 1024,
                 "%d features per lane need more LDS than a workgroup has at %d bins", c.fpl, n_bins);
  }
  set_hist_shape_config(n_feat, n_bins, n_cols, const_hessian != 0, c);
  return kOk;
}

extern "C" int wgbm_get_hist_config(int32_t n_feat, int32_t n_bins, int64_t n_cols, int32_t const_hessian,
                                    int32_t* out5) {
  const HistConfig c = fit_hist_config(n_feat, n_bins, n_cols, const_hessian != 0);
  out5[0] = c.chunk_rows;
  out5[1] = c.waves;
  out5[2] = c.fpl;
  out5[3] = c.depth;
  out5[4] = hist_slots(c);
  return kOk;
}

static int check_config(const wgbm_fit_config* c) {
  WGBM_REQUIRE(c != nullptr, "null config");
  WGBM_REQUIRE(c->n_rows >= 1 && c->n_rows * (int64_t)c->n_classes < 0x7fffffff, "row count out of range");
  WGBM_REQUIRE(c->n_feat >= 1 && c->n_cols >= 1, "feature count out of range");
  WGBM_REQUIRE(c->n_bins >= 2 && c->n_bins <= 127, "num_bins must be in [2, 127]");
  WGBM_REQUIRE(c->n_eras >= 1, "era count must be positive");
  WGBM_REQUIRE(c->n_classes >= 1 && c->n_classes <= 256, "class count out of range");
  WGBM_REQUIRE(c->max_depth >= 1 && c->max_depth <= 20, "max_depth out of range");
  WGBM_REQUIRE((int64_t)c->n_classes << c->max_depth < (1 << 24), "tree too large for level tables");
  WGBM_REQUIRE(c->objective == 0 || c->objective == 1, "unknown objective %d", c->objective);
  WGBM_REQUIRE(c->objective != 0 || c->n_classes == 1, "regression uses one channel");
  return kOk;
}

extern "C" size_t wgbm_fit_workspace_bytes(const wgbm_fit_config* c) {
  if (check_config(c) != kOk) return 0;
  return fit_layout(*c).total;
}

// One boosting round (K class trees), every launch on stream st.  Nothing
// here depends on the round index on the host: the round counter, the
// feature subset and the output offsets are read on the device, so the
// launches of one round can be captured once and replayed.
// The round's trees reach the scores by a tree walk: for regression without
// a per-round loss, at the start of the next round, in the launch that
// forms that round's gradients (apply_prev; the call's last round is walked
// on its own after the rounds, tree_walk_launch); otherwise at the end of
// the round itself.
static bool defer_tree_walk(const wgbm_fit_config& c, const wgbm_fit_io* io) {
  return c.objective == 0 && io->train_loss == nullptr;
}

// A regression tree whose levels all fuse into their tails: its partition
// records every row's leaf (a leaf's rows when the leaf is decided, the last
// split level's rows by side), so the deferred walk reads the leaves instead
// of the codes.
static bool row_leaves(const wgbm_fit_config& c, const wgbm_fit_io* io, const FitLayout& L) {
  if (!defer_tree_walk(c, io) || c.n_classes != 1 || c.max_depth < 1) return false;
  for (int d = 0; d < c.max_depth; ++d)
    if (!level_tail_fusable((int64_t)1 << d, c.n_eras, L.KN)) return false;
  return true;
}

static TreeArgs tree_args(const wgbm_fit_config& c, const FitLayout& L, const wgbm_fit_io* io, char* ws) {
  TreeArgs t{};
  t.status = io->node_status;
  t.feat = io->node_feat;
  t.bin = io->node_bin;
  t.value = io->node_value;
  t.round_ctr = reinterpret_cast<const int32_t*>(ws + L.round_ctr);
  t.K = c.n_classes;
  t.D = c.max_depth;
  t.codes_t = reinterpret_cast<const int8_t*>(ws + L.codes_t);
  t.N = c.n_rows;
  t.lr = c.learning_rate;
  t.row_leaf = row_leaves(c, io, L) ? reinterpret_cast<const int16_t*>(ws + L.row_leaf) : nullptr;
  return t;
}


static int tree_walk_launch(const wgbm_fit_config& c, const FitLayout& L, const wgbm_fit_io* io, char* ws,
                            hipStream_t st) {
  const int64_t N = c.n_rows;
  const int bpc = (int)std::min<int64_t>(ceil_div(N, 256 * kApplyRows), 4096);
  hipLaunchKernelGGL(k_apply_tree, dim3((unsigned)(c.n_classes * bpc)), dim3(256), 0, st, tree_args(c, L, io, ws),
                     io->scores, bpc);
  WGBM_LAUNCHED("k_apply_tree");
  return kOk;
}

static int enqueue_round(const wgbm_fit_config& c, const FitLayout& L, const wgbm_fit_io* io, char* ws,
                         hipStream_t st, bool apply_prev) {
  int rc = kOk;
  auto I = [&](size_t off) { return reinterpret_cast<int32_t*>(ws + off); };
  auto Fp = [&](size_t off) { return reinterpret_cast<float*>(ws + off); };
  auto G2 = [&](size_t off) { return reinterpret_cast<float2*>(ws + off); };
  const int K = c.n_classes, E = c.n_eras, D = c.max_depth;
  const int64_t N = c.n_rows;
  const int32_t* fidx = io->feat_idx ? I(L.fidx_cur) : nullptr;
    // Round start in two launches: gradients (after the previous round's
    // tree walk when it is deferred), then the root statistics with the
    // level-0 tables (rewritten every round: the partition overwrites table 0
    // at level 1) and the root histogram plan.
    RootArgs ra{};
    ra.era_count = io->era_count;
    ra.E = E;
    ra.K = K;
    ra.N = (int)N;
    ra.era_start = I(L.era_start);
    ra.seg_start = I(L.root_seg_start);
    ra.seg_count = I(L.root_seg_count);
    ra.seg_slot = I(L.root_seg_slot);
    ra.nseg = I(L.root_nseg);
    ra.t_seg_start = I(L.tab[0].seg_start);
    ra.t_seg_count = I(L.tab[0].seg_count);
    ra.t_active = I(L.tab[0].node_active);
    ra.t_node_start = I(L.tab[0].node_start);
    ra.t_node_count = I(L.tab[0].node_count);
    ra.round_ctr = I(L.round_ctr);
    ra.feat_all = io->feat_idx;
    ra.n_feat = c.n_feat;
    ra.fidx_cur = I(L.fidx_cur);
    const int nsb = stats_blocks(N);
    if (c.objective == 0) {
      hipLaunchKernelGGL(k_round_start_regression, dim3((unsigned)(nsb + 1)), dim3(256), 0, st, io->y,
                         io->scores, io->weight, N, I(L.perm_root), G2(L.gh_root), I(L.stat_part), nsb,
                         tree_args(c, L, io, ws), apply_prev ? 1 : 0, ra);
      WGBM_LAUNCHED("k_round_start_regression");
    } else {
      const int gb = (int)ceil_div(N, 256);
      hipLaunchKernelGGL(k_round_start_softmax, dim3((unsigned)(gb + 1)), dim3(256), 0, st, io->scores, io->labels,
                         io->weight, N, K, I(L.perm_root), G2(L.gh_root), gb, ra);
      WGBM_LAUNCHED("k_round_start_softmax");
      hipLaunchKernelGGL(k_gh_stats, dim3((unsigned)nsb, (unsigned)K), dim3(256), 0, st, G2(L.gh_root), N,
                         I(L.stat_part));
      WGBM_LAUNCHED("k_gh_stats");
    }
    PlanJob root_plan{};
    {
      HistArgs ra_h{};
      ra_h.n_feat = c.n_feat;
      ra_h.seg_count = I(L.root_seg_count);
      ra_h.seg_start = I(L.root_seg_start);
      ra_h.seg_slot = I(L.root_seg_slot);
      ra_h.nseg_dev = I(L.root_nseg);
      ra_h.chunk_dev = I(L.h_chunk);
      ra_h.chunk_desc = reinterpret_cast<int4*>(ws + L.h_desc);
      ra_h.chunk_total = I(L.h_total);
      ra_h.multi_desc = reinterpret_cast<int4*>(ws + L.h_mdesc);
      ra_h.merge_parts = I(L.h_mparts);
      ra_h.seg_first = I(L.h_seg_first);
      ra_h.seg_pfirst = I(L.h_seg_pfirst);
      root_plan = hist_plan_job(ra_h, L.hc, I(L.h_multi), I(L.h_n_multi));
    }
    hipLaunchKernelGGL(k_root_stats_plan, dim3((unsigned)(K + (D > 0 ? 1 : 0))), dim3(kPlanThreads), 0, st, I(L.stat_part), nsb,
                       I(L.tab[0].stats), I(L.tree_stats), K, root_plan,
                       reinterpret_cast<unsigned long long*>(ws + L.curs[0]), std::min(K, kScanLdsSegMax),
                       I(L.round_ctr));
    WGBM_LAUNCHED("k_root_stats_plan");
    for (int d = 0; d <= D; ++d) {
      const int n_nodes = K << d;
      // One era and few nodes: the level's decisions fold into its tail
      // (no decision launch, no leaf or partition plan).
      const bool fused = d < D && level_tail_fusable(n_nodes, E, L.KN);
      // after a fused last split level the leaves are recorded: nothing is left to do
      if (d == D && d > 0 && level_tail_fusable(n_nodes / 2, E, L.KN)) break;
      const FitLayout::Tables& cur = L.tab[d & 1];
      const FitLayout::Tables& nxt = L.tab[(d + 1) & 1];
      const int32_t* perm_d = d == 0 ? I(L.perm_root) : I(L.perm_pp[d & 1]);
      const float2* gh_d = d == 0 ? G2(L.gh_root) : G2(L.gh_pp[d & 1]);
      float* hg = Fp(L.hg[d & 1]);
      float* hh = Fp(L.hh[d & 1]);
      LeafArgs la{};
      la.n_nodes = n_nodes;
      la.node_active = I(cur.node_active);
      la.dec_feat = d < D ? I(L.dec_feat) : nullptr;
      la.node_start = I(cur.node_start);
      la.node_count = I(cur.node_count);
      la.seg_first = fused ? I(L.p_seg_first) : I(L.l_seg_first);
      la.nseg_dev = fused ? I(L.p_nseg) : I(L.l_nseg);
      la.multi = I(L.l_multi);
      la.n_multi = I(L.l_n_multi);
      la.perm = perm_d;
      la.gh = gh_d;
      la.n_rows = N;
      la.lam = c.lambda;
      la.lr = c.learning_rate;
      la.pred = nullptr;  // leaves are recorded; the round's trees reach the scores by the tree walk
      la.pred_stride = K;
      la.nodes_per_class = 1 << d;
      la.leaf_val = Fp(L.l_val);
      la.node_sum = nullptr;
      la.part = reinterpret_cast<I128*>(ws + L.l_part);
      la.stats = I(L.tree_stats);
      PartArgs p{};
      if (d < D) {
        p.codes = io->codes;
        p.row_stride = c.n_cols;
        p.n_rows = N;
        p.feat_idx = fidx;
        p.n_nodes = n_nodes;
        p.n_eras = E;
        p.node_active = I(cur.node_active);
        p.dec_feat = I(L.dec_feat);
        p.dec_bin = I(L.dec_bin);
        p.seg_start = I(cur.seg_start);
        p.seg_count = I(cur.seg_count);
        p.node_start = I(cur.node_start);
        p.node_count = I(cur.node_count);
        p.perm_in = perm_d;
        p.gh_in = gh_d;
        p.perm_out = I(L.perm_pp[(d + 1) & 1]);
        p.gh_out = G2(L.gh_pp[(d + 1) & 1]);
        p.flag = reinterpret_cast<uint8_t*>(ws + L.p_flag);
        p.seg_first = I(L.p_seg_first);
        p.nseg_dev = I(L.p_nseg);
        p.multi = I(L.p_multi);
        p.n_multi = I(L.p_n_multi);
        p.lbase = I(L.p_lbase);
        p.chunk_seg = I(L.p_cseg);
        p.rbase = I(L.p_rbase);
        p.cscan = I(L.p_cscan);
        p.codes_t = reinterpret_cast<const int8_t*>(ws + L.codes_t);
        p.segL = I(L.p_segL);
        p.segP = I(L.p_segP);
        p.c_seg_start = I(nxt.seg_start);
        p.c_seg_count = I(nxt.seg_count);
        p.c_active = I(nxt.node_active);
        p.c_node_start = I(nxt.node_start);
        p.c_node_count = I(nxt.node_count);
        p.c_stats = I(nxt.stats);
        p.stat_part = I(L.stat_part);
        p.b_seg_start = I(L.b_seg_start);
        p.b_seg_count = I(L.b_seg_count);
        p.b_seg_slot = I(L.b_seg_slot);
        p.n_build = I(L.n_build);
        p.sib = I(L.sib);
        p.n_split = I(L.n_split);
        p.chunk_rows = kPartChunk;
      }
      if (d < D) {
        HistArgs a{};
        a.codes = io->codes;
        a.row_stride = c.n_cols;
        a.n_rows = N;
        a.feat_idx = fidx;
        a.n_feat = c.n_feat;
        a.n_bins = c.n_bins;
        a.perm = perm_d;
        a.gh = gh_d;
        a.seg_first = I(L.h_seg_first);
        a.seg_pfirst = I(L.h_seg_pfirst);
        a.chunk_dev = I(L.h_chunk);
        a.chunk_desc = reinterpret_cast<int4*>(ws + L.h_desc);
        a.chunk_total = I(L.h_total);
        a.multi_desc = reinterpret_cast<int4*>(ws + L.h_mdesc);
        a.merge_parts = I(L.h_mparts);
        a.out_g = hg;
        a.out_h = hh;
        a.part = reinterpret_cast<longlong2*>(ws + L.h_part);
        a.pflag = I(L.h_pflag);
        a.fine = reinterpret_cast<long long*>(ws + L.h_fine);
        a.slots_per_node = E;
        a.hess_const = fit_one_plane(c) ? 1 : 0;
        a.stats = I(cur.stats);
        a.accumulate = 0;
        if (d == 0) {
          a.seg_start = I(L.root_seg_start);
          a.seg_count = I(L.root_seg_count);
          a.seg_slot = I(L.root_seg_slot);
          a.nseg_dev = I(L.root_nseg);
          a.planned = 1;  // planned beside the root statistics
          rc = hist_launch(a, L.hc, L.KN, (int64_t)K * E, I(L.h_multi), I(L.h_n_multi), st);
        } else {
          a.seg_start = I(L.b_seg_start);
          a.seg_count = I(L.b_seg_count);
          a.seg_slot = I(L.b_seg_slot);
          a.nseg_dev = I(L.n_build);
          a.planned = 1;  // planned beside the previous level's child statistics
          // each built child also writes its sibling = parent - built
          a.sub_g = Fp(L.hg[(d - 1) & 1]);
          a.sub_h = Fp(L.hh[(d - 1) & 1]);
          rc = hist_launch(a, L.hc, L.KN / 2 + 1, ((int64_t)K << (d - 1)) * E, I(L.h_multi), I(L.h_n_multi), st);
        }
        if (rc != kOk) return rc;
        LevelScanArgs s{};
        s.G = hg;
        s.H = hh;
        s.n_nodes = n_nodes;
        s.n_eras = E;
        s.n_feat = c.n_feat;
        s.n_bins = c.n_bins;
        s.node_active = I(cur.node_active);
        s.feat_idx = fidx;
        s.mono_req = io->mono_required;
        s.msg = c.min_split_gain;
        s.mcw = c.min_child_weight;
        s.lam = c.lambda;
        s.cand_agree = I(L.cand_agree);
        s.cand_crit = Fp(L.cand_crit);
        s.cand_idx = I(L.cand_idx);
        s.dec_feat = I(L.dec_feat);
        s.dec_bin = I(L.dec_bin);
        s.era_gain = Fp(L.era_gain);
        s.era_g = Fp(L.era_g);
        s.era_d = reinterpret_cast<int8_t*>(ws + L.era_d);
        s.gain_nodes = L.gain_nodes;
        if (fused) {
          rc = level_scan_launch(s, st, nullptr, nullptr, false);  // the tail picks
        } else {
          // decisions, then the leaf and partition plans, in one launch
          const PlanJob leaf_job = leaf_plan_job(la), part_job = part_plan_job(p);
          rc = level_scan_launch(s, st, &leaf_job, &part_job);
        }
        if (rc != kOk) return rc;
      }
      RecordArgs r{};
      r.d = d;
      r.D = D;
      r.K = K;
      r.E = E;
      r.n_nodes = n_nodes;
      r.active = I(cur.node_active);
      r.dec_feat = I(L.dec_feat);
      r.dec_bin = I(L.dec_bin);
      r.feat_idx = fidx;
      r.node_count = I(cur.node_count);
      r.era_gain = Fp(L.era_gain);
      r.round_ctr = I(L.round_ctr);
      r.status = io->node_status;
      r.feat = io->node_feat;
      r.bin = io->node_bin;
      r.value = io->node_value;
      r.count = io->node_count;
      r.out_era_gain = io->node_era_gain;
      if (d == 0 && D == 0 && !fused) {  // no decisions ran: the leaf plan of the only level
        const PlanJob j = leaf_plan_job(la);
        plan_launch(j.src, j.chunk_rows, j.seg_first, j.multi, j.n_multi, st);
        WGBM_LAUNCHED("k_hist_plan(leaf)");
      }
      // The plan that runs beside the child statistics: the next level's
      // histogram plan, or before the last level its leaf plan.
      PlanJob next{};
      if (d + 1 < D) {
        HistArgs na{};
        na.n_feat = c.n_feat;
        na.seg_count = I(L.b_seg_count);
        na.nseg_dev = I(L.n_build);
        na.chunk_dev = I(L.h_chunk);
        na.chunk_desc = reinterpret_cast<int4*>(ws + L.h_desc);
        na.chunk_total = I(L.h_total);
        na.multi_desc = reinterpret_cast<int4*>(ws + L.h_mdesc);
        na.merge_parts = I(L.h_mparts);
        na.seg_start = I(L.b_seg_start);
        na.seg_slot = I(L.b_seg_slot);
        na.seg_first = I(L.h_seg_first);
        na.seg_pfirst = I(L.h_seg_pfirst);
        next = hist_plan_job(na, L.hc, I(L.h_multi), I(L.h_n_multi));
      } else if (d + 1 == D && !fused) {
        LeafArgs ln = la;
        ln.n_nodes = n_nodes * 2;
        ln.node_active = I(nxt.node_active);
        ln.dec_feat = nullptr;
        ln.node_start = I(nxt.node_start);
        ln.node_count = I(nxt.node_count);
        next = leaf_plan_job(ln);
      }
      if (fused) {
        TailPick pk{};
        pk.cand_agree = I(L.cand_agree);
        pk.cand_crit = Fp(L.cand_crit);
        pk.cand_idx = I(L.cand_idx);
        pk.wg_per_node = (int)level_scan_groups(c.n_feat, E, c.n_bins);
        pk.n_bins = c.n_bins;
        pk.dec_feat = I(L.dec_feat);
        pk.dec_bin = I(L.dec_bin);
        pk.era_gain = Fp(L.era_gain);
        TailMap m{I(L.p_seg_first), I(L.p_nseg), I(L.p_cseg)};
        TailCursors tc{reinterpret_cast<unsigned long long*>(ws + L.curs[d & 1]),
                       reinterpret_cast<unsigned long long*>(ws + L.curs[(d + 1) & 1]),
                       reinterpret_cast<I128*>(ws + L.child_part),
                       row_leaves(c, io, L) ? reinterpret_cast<int16_t*>(ws + L.row_leaf) : nullptr, (1 << d) - 1};
        const bool last = d + 1 == D;
        rc = level_fused_launch(p, la, r, pk, m, tc, last, L.KN, st, last ? nullptr : &next);
      } else {
        rc = level_tail_launch(d < D ? &p : nullptr, la, r, L.KN, st, d < D ? &next : nullptr);
      }
      if (rc != kOk) return rc;
    }
    if (!defer_tree_walk(c, io)) {  // the round's trees, now recorded, added to the scores
      rc = tree_walk_launch(c, L, io, ws, st);
      if (rc != kOk) return rc;
    }
    if (io->train_loss) {  // the round's training loss, after its last score update
      hipLaunchKernelGGL(k_loss_part, dim3((unsigned)nsb), dim3(256), 0, st, io->y, c.objective == 1 ? io->labels : nullptr,
                         io->scores, io->weight, N, K, reinterpret_cast<double*>(ws + L.loss_part));
      WGBM_LAUNCHED("k_loss_part");
      hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, st, reinterpret_cast<const double*>(ws + L.loss_part), nsb,
                         I(L.round_ctr), io->train_loss);
      WGBM_LAUNCHED("k_loss_final");
    }
  return kOk;
}

__global__ void k_round_init(int32_t* round_ctr, int r0) {
  if (threadIdx.x == 0) { round_ctr[0] = r0; round_ctr[1] = r0; }
}

// Rounds replayed from one captured round need at least this many rounds.
constexpr int kGraphMinRounds = 3;
static thread_local int g_graph_rounds = 0;  // rounds of the last call replayed from a graph
static int g_graph_enabled = 1;

extern "C" int32_t wgbm_fit_graph_rounds(void) { return g_graph_rounds; }

extern "C" int wgbm_set_fit_graph(int32_t enable) {
  g_graph_enabled = enable != 0;
  return kOk;
}

// Non-blocking stream (default priority) per device for graph replays.
static hipStream_t graph_stream() {
  static hipStream_t streams[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
    streams[dev] = nullptr;
  return streams[dev];
}

extern "C" int wgbm_fit_rounds(const wgbm_fit_config* cfg, const wgbm_fit_io* io, int32_t n_rounds,
                               void* workspace, size_t workspace_bytes, hipStream_t st) {
  int rc = check_config(cfg);
  if (rc != kOk) return rc;
  const wgbm_fit_config& c = *cfg;
  const FitLayout L = fit_layout(c);
  WGBM_REQUIRE(workspace_bytes >= L.total, "fit workspace too small: need %zu bytes", L.total);
  WGBM_REQUIRE(io && io->codes && io->weight && io->scores && io->era_count, "missing fit inputs");
  WGBM_REQUIRE(c.objective == 1 ? io->labels != nullptr : io->y != nullptr, "missing targets");
  WGBM_REQUIRE(io->feat_idx != nullptr || c.n_feat <= c.n_cols, "feature subset exceeds the bin matrix");
  char* ws = static_cast<char*>(workspace);
  auto I = [&](size_t off) { return reinterpret_cast<int32_t*>(ws + off); };
  const int K = c.n_classes;
  const int64_t N = c.n_rows;

  hipLaunchKernelGGL(k_root_perm, dim3((unsigned)ceil_div(L.KN, 256)), dim3(256), 0, st, io->root_order, N, K,
                     I(L.perm_root));
  WGBM_LAUNCHED("k_root_perm");
  const bool words = c.n_cols % 4 == 0 && N % 4 == 0 && (uintptr_t)io->codes % 4 == 0;
  const bool wide = c.n_cols % 8 == 0 && N % 16 == 0 && (uintptr_t)io->codes % 8 == 0 &&
                    (uintptr_t)(ws + L.codes_t) % 16 == 0;
  if (wide) {
    const int64_t n_ctiles = ceil_div(c.n_cols, kTrTile);
    const int64_t per_xcd = ceil_div(ceil_div(N, kTrTile) * n_ctiles, 8);
    hipLaunchKernelGGL(k_transpose_codes16, dim3((unsigned)(8 * per_xcd)), dim3(256), 0, st, io->codes, N, c.n_cols,
                       reinterpret_cast<int8_t*>(ws + L.codes_t), n_ctiles, per_xcd);
    WGBM_LAUNCHED("k_transpose_codes16");
  } else if (words) {
    hipLaunchKernelGGL(k_transpose_codes4, dim3((unsigned)ceil_div(N, 64), (unsigned)ceil_div(c.n_cols, 64)),
                       dim3(256), 0, st, io->codes, N, c.n_cols, reinterpret_cast<int8_t*>(ws + L.codes_t));
    WGBM_LAUNCHED("k_transpose_codes4");
  } else {
    hipLaunchKernelGGL(k_transpose_codes, dim3((unsigned)ceil_div(N, 64), (unsigned)ceil_div(c.n_cols, 64)),
                       dim3(256), 0, st, io->codes, N, c.n_cols, reinterpret_cast<int8_t*>(ws + L.codes_t));
    WGBM_LAUNCHED("k_transpose_codes");
  }

  // Fine-row sums start at zero; every histogram read-out leaves them so
  // (set here, before any capture: a graph holds no memset).
  WGBM_HIP(hipMemsetAsync(ws + L.h_fine, 0, 8 * (size_t)L.hist_nodes * L.node_elems, st));
  g_graph_rounds = 0;
  hipLaunchKernelGGL(k_round_init, dim3(1), dim3(64), 0, st, I(L.round_ctr), 0);
  WGBM_LAUNCHED("k_round_init");
  const bool defer = defer_tree_walk(c, io);
  if (n_rounds < kGraphMinRounds || hist_timer_on() || !g_graph_enabled) {
    for (int r = 0; r < n_rounds; ++r) {
      rc = enqueue_round(c, L, io, ws, st, defer && r > 0);
      if (rc != kOk) return rc;
    }
    return defer && n_rounds > 0 ? tree_walk_launch(c, L, io, ws, st) : kOk;  // the last round's trees
  }
  // Round 0 runs eagerly (it also sets kernel attributes); round 1 is
  // captured into a graph on a side stream and replayed for rounds 1..n-1
  // (each walking the previous round's trees first when that is deferred).
  rc = enqueue_round(c, L, io, ws, st, false);
  if (rc != kOk) return rc;
  hipStream_t gs = graph_stream();
  WGBM_REQUIRE(gs != nullptr, "could not create the graph stream");
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  WGBM_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
  WGBM_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
  WGBM_HIP(hipEventRecord(ev_in, st));
  WGBM_HIP(hipStreamWaitEvent(gs, ev_in, 0));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    const int crc = enqueue_round(c, L, io, ws, gs, defer);
    const hipError_t ce = hipStreamEndCapture(gs, &graph);
    if (crc == kOk && ce == hipSuccess && graph &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      exec = nullptr;
    if (graph) (void)hipGraphDestroy(graph);
  }
  if (exec == nullptr) {
    // Capture is an optimisation: on any failure run the rounds eagerly.
    (void)hipGetLastError();
    for (int r = 1; r < n_rounds; ++r) {
      rc = enqueue_round(c, L, io, ws, gs, defer);
      if (rc != kOk) { (void)hipEventDestroy(ev_in); (void)hipEventDestroy(ev_out); return rc; }
    }
  }
  for (int r = 1; exec && r < n_rounds; ++r) {
    const hipError_t le = hipGraphLaunch(exec, gs);
    if (le != hipSuccess) { (void)hipGraphExecDestroy(exec); WGBM_HIP(le); }
    ++g_graph_rounds;
  }
  if (defer) {  // the last round's trees
    rc = tree_walk_launch(c, L, io, ws, gs);
    if (rc != kOk) {
      if (exec) (void)hipGraphExecDestroy(exec);
      (void)hipEventDestroy(ev_in);
      (void)hipEventDestroy(ev_out);
      return rc;
    }
  }
  WGBM_HIP(hipEventRecord(ev_out, gs));
  WGBM_HIP(hipStreamWaitEvent(st, ev_out, 0));
  // The executable graph must outlive its launches: wait for them here
  // (the caller reads the node tables right after this call anyway).
  WGBM_HIP(hipEventSynchronize(ev_out));
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipEventDestroy(ev_in);
  (void)hipEventDestroy(ev_out);
  return kOk;
}
