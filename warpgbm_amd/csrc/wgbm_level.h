// Level-wise tree growth: partition and leaf argument blocks.
#pragma once
#include "wgbm_stats.h"
#include "common.h"
#include "wgbm_internal.h"

namespace wgbm {

constexpr int kPartChunk = 1024;  // rows per partition workgroup (4 passes of 256)

struct PartArgs {
  const int8_t* codes;
  int64_t row_stride;
  int64_t n_rows;
  const int8_t* codes_t;        // column-major copy [columns][n_rows], or nullptr
  const int32_t* feat_idx;      // local -> global feature (nullptr: dec_feat is global)
  int32_t n_nodes;              // nodes of this level
  int32_t n_eras;
  const int32_t* node_active;   // [n_nodes]
  const int32_t* dec_feat;      // [n_nodes] local feature or -1
  const int32_t* dec_bin;       // [n_nodes]
  const int32_t* seg_start;     // [n_nodes*E]
  const int32_t* seg_count;
  const int32_t* node_start;    // [n_nodes]
  const int32_t* node_count;
  const int32_t* perm_in;       // positions of this level
  const float2* gh_in;          // or nullptr
  int32_t* perm_out;            // positions of the next level
  float2* gh_out;               // or nullptr
  // scratch
  uint8_t* flag;                // [positions] goes-left flag per position
  int32_t* seg_first;           // [n_nodes*E+1]
  int32_t* nseg_dev;            // scalar
  int32_t* multi;               // plan scratch
  int32_t* n_multi;
  int32_t* chunk_seg;           // [max chunks] segment of each chunk, written by the plan, or nullptr
  int32_t* lbase;               // [max chunks]
  int32_t* rbase;
  int32_t* cscan;               // [max chunks + 1]
  int32_t* segL;                // [n_nodes*E]
  int32_t* segP;                // [2*(n_nodes*E+1)] scratch
  // next level tables ([2*n_nodes(*E)])
  int32_t* c_seg_start;
  int32_t* c_seg_count;
  int32_t* c_active;
  int32_t* c_node_start;
  int32_t* c_node_count;
  int32_t* c_stats;             // per child: gradient statistics (nullable)
  int32_t* stat_part;           // [max chunks][2][kStatWords] scratch when c_stats is set
  // smaller-child build list ([n_nodes*E]) and sibling map ([3*n_nodes])
  int32_t* b_seg_start;
  int32_t* b_seg_count;
  int32_t* b_seg_slot;
  int32_t* n_build;
  int32_t* sib;
  int32_t* n_split;
  int32_t chunk_rows;
};

struct LeafArgs {
  int32_t n_nodes;
  const int32_t* node_active;
  const int32_t* dec_feat;      // nullptr: every active node is a leaf
  const int32_t* node_start;
  const int32_t* node_count;
  int32_t* seg_first;           // scratch [n_nodes+1]
  int32_t* nseg_dev;
  int32_t* multi;
  int32_t* n_multi;
  const int32_t* perm;
  const float2* gh;             // positional (g, h), or nullptr for gather mode
  const float* g;
  const float* h;
  int64_t n_rows;
  double lam;
  float lr;
  float* pred;                  // nullable
  int32_t pred_stride;
  int32_t nodes_per_class;
  float* leaf_val;              // [n_nodes]
  double* node_sum;             // [n_nodes*2] or nullptr
  const int32_t* stats;         // [n_nodes][kStatWords]: each leaf's max |g| and max |h| give its sum's scale
  I128* part;                   // [max chunks * 2]: exact fixed-point sums of g, h per chunk
};

// Where the tree record of one level goes (flat node arrays of the caller,
// heap order per (round, class) tree of 2^(D+1) - 1 nodes).
struct RecordArgs {
  int32_t d, D, K, E, n_nodes;
  const int32_t* active;        // [n_nodes]
  const int32_t* dec_feat;      // [n_nodes] local feature or -1 (unused at d == D)
  const int32_t* dec_bin;
  const int32_t* feat_idx;      // local -> global feature, or nullptr
  const int32_t* node_count;
  const float* era_gain;        // [n_nodes*E] winner's per-era gains (with out_era_gain)
  const int32_t* round_ctr;     // device round index
  int32_t* status;
  int32_t* feat;
  int32_t* bin;
  float* value;
  int32_t* count;
  float* out_era_gain;          // nullable
};

// The batched engine's level tail in three or four launches: partition
// counts with leaf chunk sums; the partition scan; stable scatter with leaf
// score updates and the level's tree record; child statistics with `next`
// (the next level's histogram or leaf plan) when given.  p == nullptr at
// the last level (leaves and record only).  The leaf plan must already be
// in la's plan tables.
int level_tail_launch(const PartArgs* p, const LeafArgs& la, const RecordArgs& r, int64_t rows_max, hipStream_t st,
                      const PlanJob* next);

// One-era level tail with the decisions folded in (level_tail_fusable):
// the scan's per-workgroup candidates and where the decisions go ...
struct TailPick {
  const int32_t* cand_agree;
  const float* cand_crit;
  const int32_t* cand_idx;
  int32_t wg_per_node;          // candidates per node (level_scan_groups)
  int32_t n_bins;
  int32_t* dec_feat;            // [n_nodes] local feature or -1
  int32_t* dec_bin;
  float* era_gain;              // [n_nodes] the winner's gain, or nullptr
};
// ... and the level's chunk map, written by its first launch (the leaf
// tables and the partition's chunk-to-segment map must be these arrays).
struct TailMap {
  int32_t* seg_first;           // [n_nodes + 1] first chunk of each node
  int32_t* nseg;                // scalar: n_nodes
  int32_t* chunk_node;          // [chunks] node of each chunk
};
// Per-node cursors of the single-pass split (node j's word at j *
// kCurStride: rows claimed in its left child in the high 32 bits and in its
// right child in the low 32, zero before the level) and the next level's
// (cleared for it); after the last split level, the children's exact sums
// per chunk ({left g, left h, right g, right h}).
constexpr int kCurStride = 32;  // 64-bit words: one 256-byte line per node
constexpr int kScanLdsSegMax = 512;  // most nodes of a fused level (kTailMaxNodes)
struct TailCursors {
  unsigned long long* cur;
  unsigned long long* cur_next;
  I128* child_part;
  // One class: every row of a node decided a leaf, and every row of the
  // last split level by side, gets its leaf's heap index here (null: not
  // recorded).  Node j of the level is heap node node_base + j.
  int16_t* row_leaf;
  int node_base;
};
bool level_tail_fusable(int64_t n_nodes, int n_eras, int64_t rows_max);
// A one-era level before the last in two launches: decisions with a
// single-pass split (or, last == true -- the children are leaves -- the
// children's sums without moving rows), then the child tables, statistics,
// the next level's histogram plan and the tree record (of the children too
// when last).  Leaves are recorded, not applied: the round's tree reaches the
// scores afterwards, from the row leaves tc.row_leaf records (when set) or
// by walking it over the codes.
int level_fused_launch(const PartArgs& p, const LeafArgs& la, const RecordArgs& r, const TailPick& pick,
                       const TailMap& m, const TailCursors& tc, bool last, int64_t rows_max, hipStream_t st,
                       const PlanJob* next);

// planned: the plan ran in an earlier fused launch (leaf_plan_job /
// part_plan_job beside the level's decisions).  next_hist: run the next level's histogram
// plan in the same launch as the child statistics.
int partition_launch(const PartArgs& a, int64_t rows_max, hipStream_t st, bool planned = false,
                     const PlanJob* next_hist = nullptr);
int leaf_launch(const LeafArgs& a, int64_t rows_max, hipStream_t st, bool planned = false);
PlanJob part_plan_job(const PartArgs& a);
PlanJob leaf_plan_job(const LeafArgs& a);
int64_t leaf_max_chunks(int64_t rows_max, int64_t n_nodes);
int subtract_launch(const float* pg, const float* ph, float* cg, float* ch, const int32_t* sib,
                    const int32_t* n_split, int64_t max_split, int64_t node_elems, hipStream_t st);

__global__ void k_part_single_init(int32_t* sc, int feature, int bin, int n);
__global__ void k_leaf_single_init(int32_t* sc, int n);
__global__ void k_sub_plain(const float* pg, const float* ph, const float* cg, const float* ch, float* sg,
                            float* sh, int64_t n);

}  // namespace wgbm
