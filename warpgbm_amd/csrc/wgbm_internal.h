// Internal (C++) interfaces shared between the kernel translation units.
#pragma once
#include "common.h"

namespace wgbm {

// Histogram tiling.  Sums are exact fixed-point integers, so none of these
// changes a bit of a histogram, only the speed.  chunk_rows is the largest
// chunk; each launch's plan picks its chunk size between kChunkMin and it so
// that the launch has about kHistSlots work items.
struct HistConfig {
  int chunk_rows = 0;
  int waves = 0;
  int fpl = 0;     // features per lane (load width)
  int depth = 0;   // rows whose code loads are in flight per wave (load block)
  int slots = 0;   // workgroups of a launch (work items aimed at), 0: kHistSlots
};

// Default workgroup slots of the histogram build: two per CU on MI355X's 256
// CUs.  Fewer slots mean fewer chunk partials to write and merge, more mean
// more workgroups to share the LDS-bound adds; the autotuner picks per shape.
constexpr int kHistSlots = 512;
inline int hist_slots(const HistConfig& c) { return c.slots > 0 ? c.slots : kHistSlots; }
constexpr int kChunkMin = 1024;

struct HistArgs {
  const int8_t* codes;     // [n_rows, row_stride] row-major bin codes
  int64_t row_stride;      // columns of the bin matrix
  int64_t n_rows;          // rows of the bin matrix
  const int32_t* feat_idx; // [n_feat] column of each local feature (nullptr: identity)
  int32_t n_feat;
  int32_t n_bins;
  const int32_t* perm;     // position -> row id
  const float* g;          // per row (gather mode)
  const float* h;
  const float2* gh;        // per position (positional mode), or nullptr
  const int32_t* seg_start;
  const int32_t* seg_count;
  const int32_t* seg_slot;
  int32_t* seg_first;      // [nseg+1], written by the plan kernel
  int32_t* seg_pfirst;     // [nseg] first partial slot of multi-chunk segments
  const int32_t* nseg_dev; // device scalar: number of segments
  int32_t* chunk_dev;      // device scalar: rows per chunk, written by the plan
  int4* chunk_desc;        // [max chunks] per chunk {first position, end position, slot, partial
                           // index or -1 (single-chunk segment)}, written by the plan
  int32_t* chunk_total;    // device scalar: chunks of the launch, written by the plan
  int4* multi_desc;        // [max multi] per multi-chunk segment {first partial, chunks, slot, 0}
  int32_t* merge_parts;    // device scalar: thread groups per word in the merge (1, 2, 4 or 8)
  int32_t chunk_rows;
  int32_t nslab;
  int32_t wide_ok;
  int32_t h_bits;          // per-row hessian fixed-point bits, from n_rows (hess_bits)
  float* out_g;            // [slot][n_feat][n_bins]
  float* out_h;
  longlong2* part;         // chunk partials (exact fixed-point sums of g, h)
  int32_t* pflag;          // [partial chunk][slab]: 1 when the chunk added fine rows
  // Exact fine-row gradient sums, [fine slot][n_feat][n_bins] int64, zero
  // between launches (each read-out takes its words back to zero).  The fine
  // slot of slot s is s, or with sibling derivation its parent's slot
  // (one child per parent is built).  fine_x: the fine words' extra shift.
  long long* fine;
  int32_t fine_x;
  const int32_t* stats;    // per node: max|g|, min h, max h, max|h| (see wgbm_stats.h)
  int32_t slots_per_node;  // histogram slots of one node (eras); node = slot / this
  int32_t hess_const;      // 1: the caller guarantees one hessian value for every row
  int32_t dual;            // without that guarantee: 1 launches both variants, each node built by the one
                           // its statistics select (fewer LDS adds for constant-hessian nodes, one launch
                           // more); 0 builds every node with two planes
  int32_t accumulate;      // 1: out += result (reference op), 0: out = result
  int32_t lds_pitch;       // floats between consecutive (j, bin) rows of a plane
  int32_t planned;         // 1: the plan (hist_plan_job) already ran in an earlier fused launch
  // Sibling by subtraction (level builds): with sub_g set, the result of
  // built slot s (node c = s / slots_per_node, era e) is also subtracted from
  // slot (c >> 1, e) of sub_g / sub_h (the previous level's histograms) into
  // slot (c ^ 1, e) of out: the sibling = parent - built, float32.
  const float* sub_g;
  const float* sub_h;
};

// Configuration used by every histogram build: the automatic choice with
// the process-wide override (wgbm_set_hist_config) applied.
// one_plane: the hessian is known to be constant (one LDS plane per histogram).
HistConfig pick_hist_config(int n_feat, int n_bins, int64_t row_stride, bool one_plane);
HistConfig auto_hist_config(int n_feat, int n_bins, int64_t row_stride, bool one_plane);
void set_hist_override(const HistConfig& c);
// Configuration for one shape (all-zero c removes it); precedes the override.
void set_hist_shape_config(int n_feat, int n_bins, int64_t n_cols, bool one_plane, const HistConfig& c);
size_t lds_bytes(const HistConfig& c, int n_bins, bool count_mode);
// Bounds for rows_max rows in nseg_max segments: chunks of a launch,
// multi-chunk segments, and partial words (one longlong2 each).
int64_t hist_max_chunks(const HistConfig& c, int n_feat, int64_t rows_max, int64_t nseg_max);
int64_t hist_max_multi(const HistConfig& c, int n_feat, int64_t rows_max);
size_t hist_partial_words(const HistConfig& c, int n_feat, int n_bins, int64_t rows_max);
size_t hist_partial_flags(const HistConfig& c, int n_feat, int64_t rows_max);  // [partial chunk][slab] ints
// Launch plan + build + merge.  rows_max / nseg_max bound the device-side
// row and segment counts (they size the grid).  multi/n_multi: scratch of
// hist_max_multi ints and one int; a.chunk_dev: one int.
int hist_launch(HistArgs a, const HistConfig& c, int64_t rows_max, int64_t nseg_max,
                int32_t* multi, int32_t* n_multi, hipStream_t st);
// True while wgbm_profile_begin's event timing is on (events are recorded
// around every histogram launch, so such launches are not graph-captured).
bool hist_timer_on();

// Batched level scan: candidates per (node, workgroup of 256 features),
// then a per-node pick.  Histograms are [n_nodes][E][n_feat][B].
struct LevelScanArgs {
  const float* G;
  const float* H;
  int32_t n_nodes, n_eras, n_feat, n_bins;
  const int32_t* node_active;   // [n_nodes]
  const int32_t* feat_idx;      // local -> global feature (nullptr: identity)
  const int8_t* mono_req;       // [global features] required direction, or nullptr
  float msg, mcw, lam;
  int32_t* cand_agree;          // [n_nodes * level_scan_groups(n_feat)]
  float* cand_crit;
  int32_t* cand_idx;
  int32_t* dec_feat;            // [n_nodes] local feature or -1
  int32_t* dec_bin;             // [n_nodes] bin or -1
  float* era_gain;              // [n_nodes][E] gain of the winner, or nullptr
  // E > 1: per-era gains and directions of a group of nodes
  // ([nodes of the group][E][n_feat][B-1]); level_scan_gain_nodes(...) nodes
  // fit in the buffers.
  float* era_g;
  int8_t* era_d;
  int64_t gain_nodes;           // nodes whose gains fit era_g / era_d
};
struct PlanJob;
// leaf / part: when both are given, the level's leaf and partition plans run
// after the per-node decisions (in the same single-workgroup launch for one
// era and few nodes, else in launches of their own).  pick == false (one era
// only): launch the scan alone and leave the candidates for the level tail.
int level_scan_launch(const LevelScanArgs& a, hipStream_t st, const PlanJob* leaf = nullptr,
                      const PlanJob* part = nullptr, bool pick = true);
int64_t level_scan_groups(int n_feat, int n_eras, int n_bins);  // candidate slots per node
// Nodes per pass of the E > 1 scan for a gain buffer of at most cap bytes.
int64_t level_scan_gain_nodes(int64_t n_nodes, int n_eras, int n_feat, int n_bins, size_t cap);

// Chunk plan shared by histogram, partition and leaf kernels: the row
// count of segment s is seg_count[s] (segments), node_count[s] for leaf
// nodes else -1 (leaves), or seg_count[s] for segments of split nodes else
// -1 (splits).  nseg comes from nseg_dev or, when that is null, nseg; it is
// copied to nseg_out for the kernels that follow.
enum PlanMode : int { kPlanSegments = 0, kPlanLeaves = 1, kPlanSplits = 2 };
struct PlanSrc {
  int32_t mode;
  int32_t nseg;
  const int32_t* nseg_dev;
  int32_t* nseg_out;
  const int32_t* seg_count;
  const int32_t* node_active;
  const int32_t* dec_feat;
  const int32_t* node_count;
  int32_t n_eras;
  // Adaptive chunk size (histogram plans): when chunk_out is set, the plan
  // picks rows per chunk in [chunk_min, chunk_rows] for about chunk_target
  // chunks and stores it there.
  int32_t* chunk_out;
  int32_t chunk_min, chunk_target;
  // Chunk descriptors (histogram plans): with chunk_desc set, the plan also
  // writes every chunk's {first position, end position, slot, partial index
  // or -1} from seg_start / seg_slot, and the chunk count to chunk_total.
  const int32_t* seg_start;
  const int32_t* seg_slot;
  int4* chunk_desc;
  int32_t* chunk_total;
  // With multi_desc set: {first partial, chunks, slot, 0} of every
  // multi-chunk segment, and merge_parts = 1, 2, 4 or 8 thread groups per
  // word so that each group sums at most 16 chunks where possible.
  int4* multi_desc;
  int32_t* merge_parts;
  // With chunk_seg set: the segment of every chunk (partition plans).
  int32_t* chunk_seg;
};
void plan_launch(const PlanSrc& src, int chunk_rows, int32_t* seg_first, int32_t* multi, int32_t* n_multi,
                 hipStream_t st);

// A plan to run inside a fused launch (plan_body runs on one 1024-thread
// workgroup).
struct PlanJob {
  PlanSrc src;
  int32_t chunk_rows;
  int32_t* seg_first;
  int32_t* seg_pfirst;
  int32_t* multi;
  int32_t* n_multi;
};
PlanJob hist_plan_job(const HistArgs& a, const HistConfig& c, int32_t* multi, int32_t* n_multi);
void plan_job_launch(const PlanJob& j, hipStream_t st);  // the plan in a launch of its own

}  // namespace wgbm
