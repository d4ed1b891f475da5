/*
 * warpgbm_amd C ABI -- MI355X (gfx950) engine for WarpGBM's training and
 * inference hot path.
 *
 * Every entry point takes device pointers (from any allocator: hipMalloc,
 * PyTorch's caching allocator, ...), plain sizes and the HIP stream to run
 * on.  Nothing here allocates device memory: entry points that need scratch
 * take a caller-provided workspace whose size a *_workspace_bytes query
 * returns.  Every call is asynchronous on `stream`; the return value is 0 on
 * success, otherwise an error code with a message in wgbm_last_error().
 *
 * The first four "reference ops" replace the pybind11 functions of the
 * reference's extension module warpgbm.cuda.node_kernel
 * (warpgbm/cuda/node_kernel.cpp:36-41) with the same argument meaning.
 * The remaining entry points are this engine's own (per-node helpers of the
 * plain path, the batched level-wise trainer and the serving predictor).
 */
#ifndef WARPGBM_AMD_H
#define WARPGBM_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- library */
const char* wgbm_last_error(void);
int wgbm_abi_version(void);
const char* wgbm_build_arch(void);
/* 0 if the current device is gfx950; its arch name is copied to name_out. */
int wgbm_check_device(char* name_out, int32_t name_len);

/* ---------------------------------------------------------------- reference ops */

/* Replaces compute_histogram3 (warpgbm/cuda/node_kernel.cpp:37, kernel
 * warpgbm/cuda/histogram_kernel.cu:5-107).  For every local feature j and
 * listed row r = sample_indices[i]: if 0 <= code < num_bins then
 * grad_hist[era[r], j, code] += grads[r] and hess_hist[...] += hess[r].
 * Histograms are [n_eras, n_feat, num_bins] float32 and ACCUMULATE (the
 * caller zero-fills).  Deterministic: see the histogram kernel notes.
 * The reference's threads_per_block / rows_per_thread launch knobs are not
 * arguments: this engine's tiling comes from wgbm_set_hist_config. */
/* Workspace of wgbm_compute_histogram3 for n_samples listed rows, n_feat
 * local features of a bin matrix with n_master_feat columns (the column
 * count decides which tilings are possible). */
size_t wgbm_histogram3_workspace_bytes(int64_t n_samples, int32_t n_feat, int32_t num_bins, int32_t n_eras,
                                       int64_t n_master_feat);
int wgbm_compute_histogram3(const int8_t* bin_indices, int64_t n_total_rows, int64_t n_master_feat,
                            const float* grads, const float* hess, const int32_t* sample_indices,
                            int64_t n_samples, const int32_t* feature_indices, int32_t n_feat,
                            const int32_t* era_indices, float* grad_hist, float* hess_hist, int32_t n_eras,
                            int32_t num_bins, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Replaces compute_split (warpgbm/cuda/node_kernel.cpp:38, kernel
 * warpgbm/cuda/best_split_kernel.cu:5-88): per-era gain and direction
 * [n_eras, n_feat, num_bins-1] float32 from histograms [n_eras, n_feat,
 * num_bins].  min_split_gain is accepted and unused, as in the reference. */
int wgbm_compute_split(const float* grad_hist, const float* hess_hist, int32_t n_eras, int32_t n_feat,
                       int32_t num_bins, float min_split_gain, float min_child_weight, float lambda,
                       float* per_era_gain, float* per_era_direction, hipStream_t stream);

/* Replaces custom_cuda_binner (warpgbm/cuda/node_kernel.cpp:39, kernel
 * warpgbm/cuda/binner.cu:5-52): out[i] = number of leading edges with
 * x[i] >= edge (NaN -> 0), one float32 column of n values. */
int wgbm_bin_column(const float* x, int64_t n, const float* bin_edges, int32_t n_edges, int8_t* out,
                    hipStream_t stream);

/* Replaces predict_forest (warpgbm/cuda/node_kernel.cpp:40, kernel
 * warpgbm/cuda/predict.cu:5-76): out[i] += sum_t lr * leaf_t(row i) over a
 * forest flattened to [n_trees, max_nodes, 6] float32 rows
 * (feature, bin, left, right, is_leaf, value).  Trees are added in tree
 * order (deterministic). */
int wgbm_predict_forest(const int8_t* bin_indices, int64_t n_rows, int64_t n_feat, const float* trees,
                        int64_t n_trees, int64_t max_nodes, float learning_rate, float* out, hipStream_t stream);

/* ---------------------------------------------------------------- plain per-node path */

/* Best split of one node from compute_split's outputs with the reference's
 * selection (warpgbm/core.py:1322-1362): out_feat_bin[0..1] = (local
 * feature, bin) or (-1, -1).  mono_required: required direction per global
 * feature (int8, -1/0/+1) or NULL; feature_indices maps local -> global
 * (NULL: identity). */
int wgbm_select_split(const float* per_era_gain, const float* per_era_direction, int32_t n_eras, int32_t n_feat,
                      int32_t num_bins, float min_split_gain, const int8_t* mono_required,
                      const int32_t* feature_indices, int32_t* out_feat_bin, hipStream_t stream);

/* Stable split of a node's row list by code[row, feature] <= bin
 * (warpgbm/core.py:1425-1427): out_rows = left rows then right rows, each
 * in input order; *n_left_out (device int32) = number of left rows. */
size_t wgbm_partition_workspace_bytes(int64_t n_samples);
int wgbm_partition_rows(const int8_t* bin_indices, int64_t n_total_rows, int64_t n_master_feat,
                        const int32_t* rows, int64_t n, int32_t feature, int32_t bin, int32_t* out_rows,
                        int32_t* n_left_out, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Leaf value sum(g)/(sum(h)+lambda) over a node's rows (warpgbm/core.py:
 * 1377-1380; the sums exact -- each row's g and h as a 62-bit fixed-point
 * integer summed in 128 bits, so the order of the rows does not matter --
 * scaled by scale_stats, the 4-int statistics wgbm_row_stats gives for the
 * leaf's tree, or with NULL by the leaf's own rows) to *leaf_out; if pred is
 * not NULL also pred[row*pred_stride + pred_col] += f32(lr) * leaf for each
 * row (warpgbm/core.py:1381-1386). */
size_t wgbm_leaf_workspace_bytes(int64_t n_samples);
int wgbm_leaf_value(const float* grads, const float* hess, const int32_t* rows, int64_t n, int64_t n_total_rows,
                    double lambda, float learning_rate, float* pred, int32_t pred_stride, int32_t pred_col,
                    float* leaf_out, const int32_t* scale_stats, void* workspace, size_t workspace_bytes,
                    hipStream_t stream);
/* Statistics of a row list -- max |g|, min h, max h, max |h| as
 * order-preserving ints -- into out_stats (4 device ints); a tree's root rows
 * give the scale of its leaves' sums (wgbm_leaf_value).  No counterpart in
 * the reference: the exact leaf sums need it. */
size_t wgbm_row_stats_workspace_bytes(void);
int wgbm_row_stats(const float* grads, const float* hess, const int32_t* rows, int64_t n, int64_t n_total_rows,
                   int32_t* out_stats, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* sibling = parent - child, float32 elementwise (warpgbm/core.py:1445-1455). */
int wgbm_hist_subtract(const float* parent_g, const float* parent_h, const float* child_g, const float* child_h,
                       float* sib_g, float* sib_h, int64_t n, hipStream_t stream);

/* Regression targets g = (y - pred) * w, h = w (warpgbm/core.py:1545-1547). */
int wgbm_grad_regression(const float* y, const float* pred, const float* w, int64_t n, float* g_out, float* h_out,
                         hipStream_t stream);

/* Multi-output targets with shared splits (warpgbm/core.py:1666-1679), y,
 * pred, w [n, n_outputs] row-major; per output m, valid = isfinite(y):
 * g[m*n + i] = valid ? (y - pred) * w : 0, h[m*n + i] = valid ? w : 0, and
 * split_g[i] / split_h[i] = float32 sums of g / h over m in output order. */
int wgbm_grad_multi_regression(const float* y, const float* pred, const float* w, int64_t n, int32_t n_outputs,
                               float* g_out, float* h_out, float* split_g, float* split_h, hipStream_t stream);

/* Softmax targets for all classes (warpgbm/core.py:339-367, :1755-1756):
 * g[k*n + i] = -(p_ik - [y_i == k]) * w_i, h[k*n + i] = max(p_ik(1-p_ik), 1e-6) * w_i. */
int wgbm_grad_softmax(const float* scores, const int32_t* labels, const float* w, int64_t n, int32_t n_classes,
                      float* g_out, float* h_out, hipStream_t stream);

/* ---------------------------------------------------------------- binning */

/* Quantile edges from column-sorted data: sorted [n, ncol] float32 (each
 * column ascending, NaN last) -> edges[c*edge_row_stride + j-1], j = 1..B-1,
 * with torch.quantile's float32 linear interpolation at
 * torch.linspace(0, 1, B+1)[1:-1] (warpgbm/core.py:1271-1278). */
int wgbm_edges_from_sorted(const float* sorted, int64_t n, int32_t ncol, int32_t num_bins, float* edges,
                           int32_t edge_row_stride, hipStream_t stream);

/* Quantile edges of every column of a row-major float32 matrix X [n_rows,
 * ld] (first n_cols columns) without sorting: a radix select of the order
 * statistics each edge needs, then torch.quantile's linear interpolation.
 * edges [n_cols][edge_row_stride], num_bins-1 used per column.  Same values
 * as wgbm_edges_from_sorted on the sorted columns (zeros order -0.0 first).
 * Replaces the per-feature torch.quantile loop of preprocess_gpu_data
 * (warpgbm/core.py:1271-1278). */
size_t wgbm_quantile_edges_workspace_bytes(int64_t n_rows, int32_t n_cols);
int wgbm_quantile_edges(const float* X, int64_t n_rows, int32_t n_cols, int64_t ld, int32_t num_bins, float* edges,
                        int32_t edge_row_stride, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Whole-matrix binning: X [n, n_feat] row-major float32 with edges
 * [n_feat, n_edges] -> codes [n, n_feat] int8, one launch
 * (replaces the per-column loops of warpgbm/core.py:1271-1282, :1893-1898). */
size_t wgbm_bin_matrix_workspace_bytes(int32_t n_feat, int32_t n_edges);
int wgbm_bin_matrix(const float* X, int64_t n, int32_t n_feat, const float* edges, int32_t n_edges, int8_t* out,
                    void* workspace, size_t workspace_bytes, hipStream_t stream);

/* ---------------------------------------------------------------- batched level-wise training */

/* Histogram tiling shared by every histogram build (plain and batched):
 * rows per chunk, waves per workgroup, features per lane (load width), rows
 * whose code loads are in flight per wave, and workgroup slots per launch
 * (work items aimed at: 128, 256, 512 or 1024; fewer slots, fewer chunk
 * partials to merge).  0 = automatic.  Sums are exact fixed-point integers,
 * so the tiling changes speed, never a bit of a histogram.
 * wgbm_get_hist_config returns the configuration used for a shape
 * (const_hessian: every row has the same hessian, as in a regression fit
 * with constant weights): out5 = {chunk_rows, waves, fpl, depth, slots}. */
int wgbm_set_hist_config(int32_t chunk_rows, int32_t waves, int32_t features_per_lane, int32_t rows_in_flight,
                         int32_t slots);
int wgbm_get_hist_config(int32_t n_feat, int32_t num_bins, int64_t n_cols, int32_t const_hessian, int32_t* out5);
/* The same knobs for one histogram shape only (n_feat features of a bin
 * matrix with n_cols columns, num_bins bins, const_hessian as above); they
 * take precedence over wgbm_set_hist_config for that shape.  All-zero knobs
 * remove the entry.  The autotuner installs its choice here, so a tiling
 * tuned for one shape never reaches another.  A configuration that does not
 * fit a workgroup's LDS at some shape is replaced by the automatic one.
 * Takes over from the reference's threads_per_block / rows_per_thread
 * launch knobs (warpgbm/core.py:68-69, :200-207). */
int wgbm_set_hist_config_shape(int32_t n_feat, int32_t num_bins, int64_t n_cols, int32_t const_hessian,
                               int32_t chunk_rows, int32_t waves, int32_t features_per_lane, int32_t rows_in_flight,
                               int32_t slots);

/* Live timing of histogram launches with HIP events on their streams:
 * begin clears and enables; end waits and returns
 * out4 = {build ms, build launches, merge ms, merge launches}. */
int wgbm_profile_begin(void);
int wgbm_profile_end(double* out4);

typedef struct wgbm_fit_config {
  int64_t n_rows;           /* N */
  int64_t n_cols;           /* columns of the bin matrix */
  int32_t n_feat;           /* features considered per tree (colsample) */
  int32_t n_bins;           /* B */
  int32_t n_eras;           /* E */
  int32_t n_classes;        /* K class channels (1 for regression) */
  int32_t max_depth;        /* D */
  int32_t objective;        /* 0 regression, 1 softmax */
  float learning_rate;
  float min_split_gain;
  float min_child_weight;
  float lambda;             /* L2_reg */
  int32_t const_weight;     /* 1: every weight[] entry holds the same value (regression
                               hessians are then constant and histograms need one LDS
                               plane); 0: unknown */
} wgbm_fit_config;

typedef struct wgbm_fit_io {
  const int8_t* codes;          /* [n_rows, n_cols] */
  const float* y;               /* [n_rows] (regression) */
  const int32_t* labels;        /* [n_rows] (softmax) */
  const float* weight;          /* [n_rows] */
  const int32_t* root_order;    /* [n_rows] rows grouped by era (stable); NULL = identity */
  const int32_t* era_count;     /* [n_eras] rows of each era in root_order */
  const int8_t* mono_required;  /* [n_cols] required split direction, or NULL */
  const int32_t* feat_idx;      /* [n_rounds, n_feat] features per round, or NULL (identity) */
  float* scores;                /* [n_rows, n_classes] running scores, updated in place */
  /* per round, per class, per heap node (2^(D+1)-1), level order: */
  int32_t* node_status;         /* 0 absent, 1 split, 2 leaf */
  int32_t* node_feat;           /* global feature of a split */
  int32_t* node_bin;            /* split bin */
  float* node_value;            /* leaf value */
  int32_t* node_count;          /* rows in the node */
  float* node_era_gain;         /* [.., n_eras] per-era gain of a split, or NULL */
  /* [n_rounds] training loss after each round, or NULL: weighted MSE of
   * the scores (regression) or weighted log loss of their softmax, clamped
   * to [1e-15, 1 - 1e-15] (multiclass) -- the per-tree training loss that
   * compute_eval / compute_eval_multiclass log (warpgbm/core.py:1488-1516,
   * :1793-1845; metrics.py weighted_mse_torch / weighted_log_loss_torch),
   * summed in float64 here. */
  float* train_loss;
} wgbm_fit_io;

size_t wgbm_fit_workspace_bytes(const wgbm_fit_config* config);
/* Trains n_rounds boosting rounds (n_rounds * n_classes trees) without any
 * host synchronisation.  Replaces grow_forest / grow_forest_multiclass
 * (warpgbm/core.py:1518-1580, :1719-1791) and their per-node recursion
 * grow_tree (:1364-1470). */
int wgbm_fit_rounds(const wgbm_fit_config* config, const wgbm_fit_io* io, int32_t n_rounds, void* workspace,
                    size_t workspace_bytes, hipStream_t stream);
/* Rounds of the last wgbm_fit_rounds call on this thread that were replayed
 * from a captured graph (one round is captured once, then relaunched; 0 when
 * the call ran eagerly: fewer than 3 rounds, or histogram timing on).
 * Diagnostic; no counterpart in the reference. */
int32_t wgbm_fit_graph_rounds(void);
/* Process-wide switch for that capture (default on); off, every round is
 * launched eagerly (e.g. under per-dispatch counter collection).  No
 * counterpart in the reference. */
int wgbm_set_fit_graph(int32_t enable);

/* ---------------------------------------------------------------- serving predictor */

/* Compile level-order node tables (as written by wgbm_fit_rounds, src_nodes
 * heap slots per tree) into a complete heap forest of the given depth:
 * nodes [n_trees, 2^depth - 1] uint32 (feature << 8 | bin), leaves
 * [n_trees, 2^depth] float32 pre-multiplied by the learning rate. */
int wgbm_heap_compile(const int32_t* status, const int32_t* feat, const int32_t* bin, const float* value,
                      int32_t src_nodes, int32_t n_trees, int32_t depth, float learning_rate, uint32_t* nodes,
                      float* leaves, hipStream_t stream);

/* out[row*K + k] = base[k] + sum over class-k trees; class k owns trees
 * [class_first[k], class_first[k+1]) in blocks of tree_block trees. */
size_t wgbm_heap_predict_workspace_bytes(int64_t n_rows, int32_t n_classes, int32_t n_blocks);
int wgbm_heap_predict(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                      const float* leaves, int32_t n_trees, int32_t depth, int32_t tree_block,
                      const int32_t* class_first, const int32_t* blocks_of_class, int32_t n_classes,
                      int32_t n_blocks, const float* base, float* out, void* workspace, size_t workspace_bytes,
                      hipStream_t stream);
/* Heap-forest predict in tree order: out[row*K + k] = base[k] + lr * leaf of
 * every class-k tree, added one tree at a time in tree order (no
 * workspace).  The scores of a row equal, bit for bit, the scores training
 * accumulated for it and what wgbm_predict_forest adds on the reference's
 * flattened forest; this is the estimator's predict path
 * (warpgbm/core.py:1902-1962, :2021-2058) with a compiled forest instead of
 * the per-tree host flattening of core.py:2283-2328. */
int wgbm_heap_predict_tree_order(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                                 const float* leaves, int32_t n_trees, int32_t depth, const int32_t* class_first,
                                 int32_t n_classes, const float* base, float* out, hipStream_t stream);
/* wgbm_heap_predict with no workspace and no allocation of any kind.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The per-block partial sums live in a
 * file-scope device array of 4 Mi floats (16 MiB) that every GPU the library
 * runs on holds until the process ends; calls with n_rows * n_classes *
 * n_blocks above 4 Mi return an error.  Every call on one GPU shares that
 * GPU's array, so calls that overlap in time on one GPU race: serialise
 * them (one stream, or events).  Results equal wgbm_heap_predict's bit for
 * bit. */
int wgbm_heap_predict_noalloc(const int8_t* codes, int64_t n_rows, int32_t n_feat, const uint32_t* nodes,
                              const float* leaves, int32_t n_trees, int32_t depth, int32_t tree_block,
                              const int32_t* class_first, const int32_t* blocks_of_class, int32_t n_classes,
                              int32_t n_blocks, const float* base, float* out, hipStream_t stream);
/* Traversal variant of wgbm_heap_predict, for tuning (process-wide): 0
 * automatic, 1 node reads through the cache with two trees per lane in
 * flight, 2 the same with four, 3 nodes staged in LDS 16 trees at a time.
 * Every variant adds a row's trees in the same order: results are equal bit
 * for bit.  No counterpart in the reference. */
int wgbm_set_predict_mode(int32_t mode);

#ifdef __cplusplus
}
#endif
#endif /* WARPGBM_AMD_H */
