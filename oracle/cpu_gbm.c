/* Multithreaded FP64 C restatement of the WarpGBM tree growth.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py and a checker
 * in tests/.  It restates the NumPy oracle in gbm_oracle.py (each function
 * names the reference lines it follows) with OpenMP threads, and is pinned
 * against that oracle by tests/test_oracle.py (same trees, leaves within
 * 1e-9 relative).  The product path never loads it.
 *
 * Arrays: codes int8 [n_rows][n_cols] row-major; per-row float64 targets;
 * trees as heap tables of 2^(D+1) - 1 nodes (root 0, children 2i+1 / 2i+2):
 * status 0 absent / 1 split / 2 leaf, feature, bin, leaf value, row count.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  const int8_t* codes;
  int64_t n_rows;
  int32_t F;  /* columns = features */
  const double* g;
  const double* h;
  const int32_t* era; /* NULL: one era */
  int32_t E, B, D;
  double mcw, msg, lam;
  const int8_t* mono; /* [F] required direction, or NULL */
  int32_t threads;
  int32_t* pos;  /* positions -> row, grouped by node */
  int32_t* tmp;  /* partition scratch */
  double* gain;  /* [E][F][B-1] */
  double* dir;
  int32_t *status, *feat, *bin, *count;
  double* value;
  int32_t* leaf_of_row; /* or NULL */
} Ctx;

static size_t hist_len(const Ctx* x) { return (size_t)x->E * x->F * x->B; }

/* G, H interleaved [E][F][B][2] over positions [s, s + c): the histogram of
 * warpgbm/cuda/histogram_kernel.cu:33-48 (codes outside [0, B) dropped),
 * rows split over threads with private sums added in thread order. */
static void build_hist(const Ctx* x, int64_t s, int64_t c, double* GH) {
  const size_t L = hist_len(x);
  const int64_t F = x->F;
  const int B = x->B;
  int T = x->threads;
  if (c * F < (int64_t)1 << 18) T = 1;
  if (T <= 1) {
    memset(GH, 0, 2 * L * sizeof(double));
    for (int64_t i = s; i < s + c; ++i) {
      const int32_t r = x->pos[i];
      const int8_t* cr = x->codes + (int64_t)r * F;
      const double gr = x->g[r], hr = x->h[r];
      double* base = GH + 2 * (size_t)(x->era ? x->era[r] : 0) * F * B;
      for (int64_t f = 0; f < F; ++f) {
        const int v = cr[f];
        if ((unsigned)v < (unsigned)B) {
          base[2 * (f * B + v)] += gr;
          base[2 * (f * B + v) + 1] += hr;
        }
      }
    }
    return;
  }
  double* part = (double*)malloc((size_t)T * 2 * L * sizeof(double));
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    double* P = part + (size_t)t * 2 * L;
    memset(P, 0, 2 * L * sizeof(double));
    const int64_t i0 = s + c * t / T, i1 = s + c * (t + 1) / T;
    for (int64_t i = i0; i < i1; ++i) {
      const int32_t r = x->pos[i];
      const int8_t* cr = x->codes + (int64_t)r * F;
      const double gr = x->g[r], hr = x->h[r];
      double* base = P + 2 * (size_t)(x->era ? x->era[r] : 0) * F * B;
      for (int64_t f = 0; f < F; ++f) {
        const int v = cr[f];
        if ((unsigned)v < (unsigned)B) {
          base[2 * (f * B + v)] += gr;
          base[2 * (f * B + v) + 1] += hr;
        }
      }
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t k = 0; k < (int64_t)(2 * L); ++k) {
      double a = 0.0;
      for (int u = 0; u < T; ++u) a += part[(size_t)u * 2 * L + k];
      GH[k] = a;
    }
  }
  free(part);
}

/* Per-era gain and direction (warpgbm/cuda/best_split_kernel.cu:25-54):
 * prefix sums over bins b < B-1; both sides need H >= mcw. */
static void split_gains(const Ctx* x, const double* GH) {
  const int B = x->B;
  const int64_t EF = (int64_t)x->E * x->F;
#pragma omp parallel for num_threads(x->threads) schedule(static) if (EF * B > 65536)
  for (int64_t ef = 0; ef < EF; ++ef) {
    const double* p = GH + 2 * ef * B;
    double Gt = 0.0, Ht = 0.0;
    for (int b = 0; b < B; ++b) { Gt += p[2 * b]; Ht += p[2 * b + 1]; }
    double GL = 0.0, HL = 0.0;
    double* go = x->gain + ef * (B - 1);
    double* d = x->dir + ef * (B - 1);
    for (int b = 0; b < B - 1; ++b) {
      GL += p[2 * b];
      HL += p[2 * b + 1];
      const double GR = Gt - GL, HR = Ht - HL;
      if (HL >= x->mcw && HR >= x->mcw) {
        go[b] = GL * GL / (HL + x->lam) + GR * GR / (HR + x->lam) - Gt * Gt / (Ht + x->lam);
        d[b] = (GL / (HL + x->lam) > GR / (HR + x->lam)) ? 1.0 : -1.0;
      } else {
        go[b] = 0.0;
        d[b] = 0.0;
      }
    }
  }
}

/* Directional era selection of one node (warpgbm/core.py:1322-1362):
 * E == 1: crit = gain, candidates crit > msg; E > 1: crit = mean gain over
 * eras, only candidates whose |sum of directions| equals the maximum over
 * every (feature, bin); a monotone feature needs the required direction
 * (in every era).  First index of the largest crit, or -1. */
static int select_split(const Ctx* x, int* fl_out, int* bin_out) {
  const int Bm1 = x->B - 1;
  const int64_t FB = (int64_t)x->F * Bm1;
  const int E = x->E;
  int maxag = 0;
  if (E > 1) {
    for (int64_t k = 0; k < FB; ++k) {
      double sd = 0.0;
      for (int e = 0; e < E; ++e) sd += x->dir[e * FB + k];
      const int ag = abs((int)llrint(sd));
      if (ag > maxag) maxag = ag;
    }
  }
  int64_t best = -1;
  double bc = -INFINITY;
  for (int64_t k = 0; k < FB; ++k) {
    const int f = (int)(k / Bm1);
    const int req = x->mono ? x->mono[f] : 0;
    double crit;
    int ok;
    if (E == 1) {
      crit = x->gain[k];
      ok = crit > x->msg && (req == 0 || x->dir[k] == (double)req);
    } else {
      double sg = 0.0, sd = 0.0;
      int mono_ok = 1;
      for (int e = 0; e < E; ++e) {
        sg += x->gain[e * FB + k];
        sd += x->dir[e * FB + k];
        if (req != 0 && x->dir[e * FB + k] != (double)req) mono_ok = 0;
      }
      crit = sg / E;
      ok = abs((int)llrint(sd)) == maxag && crit > x->msg && mono_ok;
    }
    if (ok && crit > bc) {
      bc = crit;
      best = k;
    }
  }
  if (best < 0) return 0;
  *fl_out = (int)(best / Bm1);
  *bin_out = (int)(best % Bm1);
  return 1;
}

/* Leaf value sum(g) / (sum(h) + lambda) over the node's rows, in position
 * order (warpgbm/core.py:1377-1380). */
static void make_leaf(Ctx* x, int64_t s, int64_t c, int heap) {
  double sg = 0.0, sh = 0.0;
  for (int64_t i = s; i < s + c; ++i) {
    sg += x->g[x->pos[i]];
    sh += x->h[x->pos[i]];
  }
  x->status[heap] = 2;
  x->feat[heap] = -1;
  x->bin[heap] = -1;
  x->value[heap] = sg / (sh + x->lam);
  x->count[heap] = (int32_t)c;
  if (x->leaf_of_row)
    for (int64_t i = s; i < s + c; ++i) x->leaf_of_row[x->pos[i]] = heap;
}

/* Stable split of positions [s, s + c): code <= bin first, each side in
 * order (warpgbm/core.py:1425-1427).  Returns the left count. */
static int64_t partition(Ctx* x, int64_t s, int64_t c, int f, int b) {
  int64_t nl = 0, nr = 0;
  for (int64_t i = s; i < s + c; ++i) {
    const int32_t r = x->pos[i];
    if ((int)x->codes[(int64_t)r * x->F + f] <= b) x->pos[s + nl++] = r;
    else x->tmp[nr++] = r;
  }
  memcpy(x->pos + s + nl, x->tmp, (size_t)nr * sizeof(int32_t));
  return nl;
}

/* Depth-first growth (warpgbm/core.py:1364-1470): GH is this node's
 * histogram (its buffer is then reused for a child); the smaller child is
 * built (left on ties), the other is parent - built.  Children at the last
 * depth are leaves and take no histogram. */
static void grow(Ctx* x, double* GH, int64_t s, int64_t c, int depth, int heap) {
  int fl, b;
  if (depth == x->D) { make_leaf(x, s, c, heap); return; }
  split_gains(x, GH);
  if (!select_split(x, &fl, &b)) { make_leaf(x, s, c, heap); return; }
  x->status[heap] = 1;
  x->feat[heap] = fl;
  x->bin[heap] = b;
  x->value[heap] = 0.0;
  x->count[heap] = (int32_t)c;
  const int64_t nl = partition(x, s, c, fl, b), nr = c - nl;
  const int lh = 2 * heap + 1, rh = 2 * heap + 2;
  if (depth + 1 == x->D) {
    make_leaf(x, s, nl, lh);
    make_leaf(x, s + nl, nr, rh);
    return;
  }
  const size_t L2 = 2 * hist_len(x);
  double* GC = (double*)malloc(L2 * sizeof(double));
  const int left_built = nl <= nr;
  if (left_built) build_hist(x, s, nl, GC);
  else build_hist(x, s + nl, nr, GC);
  for (size_t k = 0; k < L2; ++k) GH[k] -= GC[k];  /* GH: the other child */
  if (left_built) {
    grow(x, GC, s, nl, depth + 1, lh);
    grow(x, GH, s + nl, nr, depth + 1, rh);
  } else {
    grow(x, GH, s, nl, depth + 1, lh);
    grow(x, GC, s + nl, nr, depth + 1, rh);
  }
  free(GC);
}

static int heap_nodes(int D) { return (2 << D) - 1; }

/* One tree on all rows (positions 0 .. n-1 in row order). */
static int grow_one(Ctx* x) {
  const int NH = heap_nodes(x->D);
  for (int i = 0; i < NH; ++i) { x->status[i] = 0; x->feat[i] = -1; x->bin[i] = -1; x->value[i] = 0.0; x->count[i] = 0; }
  for (int64_t i = 0; i < x->n_rows; ++i) x->pos[i] = (int32_t)i;
  double* GH = (double*)malloc(2 * hist_len(x) * sizeof(double));
  if (!GH) return -1;
  build_hist(x, 0, x->n_rows, GH);
  grow(x, GH, 0, x->n_rows, 0, 0);
  free(GH);
  return 0;
}

static int ctx_init(Ctx* x, const int8_t* codes, int64_t n, int32_t F, const int32_t* era, int32_t E, int32_t B,
                    int32_t D, double mcw, double msg, double lam, const int8_t* mono, int32_t threads) {
  memset(x, 0, sizeof(*x));
  if (n < 1 || F < 1 || E < 1 || B < 2 || D < 0 || D > 24) return -1;
  x->codes = codes; x->n_rows = n; x->F = F; x->era = era; x->E = E; x->B = B; x->D = D;
  x->mcw = mcw; x->msg = msg; x->lam = lam; x->mono = mono;
  x->threads = threads > 0 ? threads : omp_get_max_threads();
  x->pos = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  x->tmp = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  x->gain = (double*)malloc((size_t)E * F * (B - 1) * sizeof(double));
  x->dir = (double*)malloc((size_t)E * F * (B - 1) * sizeof(double));
  return (x->pos && x->tmp && x->gain && x->dir) ? 0 : -1;
}

static void ctx_free(Ctx* x) { free(x->pos); free(x->tmp); free(x->gain); free(x->dir); }

/* One tree from given targets g, h (float64 per row). */
int gbm_cpu_grow_tree(const int8_t* codes, int64_t n, int32_t F, const double* g, const double* h,
                      const int32_t* era, int32_t E, int32_t B, int32_t D, double mcw, double msg, double lam,
                      const int8_t* mono, int32_t threads, int32_t* status, int32_t* feat, int32_t* bin,
                      double* value, int32_t* count, int32_t* leaf_of_row) {
  Ctx x;
  if (ctx_init(&x, codes, n, F, era, E, B, D, mcw, msg, lam, mono, threads)) { ctx_free(&x); return -1; }
  x.g = g; x.h = h;
  x.status = status; x.feat = feat; x.bin = bin; x.value = value; x.count = count; x.leaf_of_row = leaf_of_row;
  const int rc = grow_one(&x);
  ctx_free(&x);
  return rc;
}

/* Regression boosting (warpgbm/core.py:1518-1580, base :926): base = mean
 * of y (unweighted), each tree fitted to g = (y - F) w, h = w, every row's
 * score then += lr * its leaf.  Tables [n_trees][2^(D+1) - 1]; scores [n]. */
int gbm_cpu_fit_regression(const int8_t* codes, int64_t n, int32_t F, const double* y, const double* w,
                           const int32_t* era, int32_t E, int32_t B, int32_t D, int32_t n_trees, double lr,
                           double mcw, double msg, double lam, const int8_t* mono, int32_t threads, double* base_out,
                           int32_t* status, int32_t* feat, int32_t* bin, double* value, int32_t* count,
                           double* scores) {
  Ctx x;
  if (ctx_init(&x, codes, n, F, era, E, B, D, mcw, msg, lam, mono, threads)) { ctx_free(&x); return -1; }
  double* g = (double*)malloc((size_t)n * sizeof(double));
  double* h = (double*)malloc((size_t)n * sizeof(double));
  int32_t* leaf = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  if (!g || !h || !leaf) { free(g); free(h); free(leaf); ctx_free(&x); return -1; }
  double base = 0.0;
  for (int64_t i = 0; i < n; ++i) base += y[i];
  base /= (double)n;
  *base_out = base;
  for (int64_t i = 0; i < n; ++i) scores[i] = base;
  const int NH = heap_nodes(D);
  int rc = 0;
  x.g = g; x.h = h; x.leaf_of_row = leaf;
  for (int t = 0; t < n_trees && rc == 0; ++t) {
    for (int64_t i = 0; i < n; ++i) { g[i] = (y[i] - scores[i]) * w[i]; h[i] = w[i]; }
    x.status = status + (int64_t)t * NH; x.feat = feat + (int64_t)t * NH; x.bin = bin + (int64_t)t * NH;
    x.value = value + (int64_t)t * NH; x.count = count + (int64_t)t * NH;
    rc = grow_one(&x);
    for (int64_t i = 0; i < n; ++i) scores[i] += lr * x.value[leaf[i]];
  }
  free(g); free(h); free(leaf);
  ctx_free(&x);
  return rc;
}

/* Softmax boosting rounds (warpgbm/core.py:1719-1791, targets :339-367):
 * scores [n][K] start at log(count_k / n + 1e-10); a round takes
 * p = softmax(scores), g_k = -(p_k - [y == k]) w, h_k = max(p_k (1 - p_k),
 * 1e-6) w, grows the K class trees and adds lr * leaf to column k.
 * Tables [n_rounds][K][2^(D+1) - 1]. */
int gbm_cpu_fit_softmax(const int8_t* codes, int64_t n, int32_t F, const int32_t* labels, int32_t K,
                        const double* w, int32_t B, int32_t D, int32_t n_rounds, double lr, double mcw, double msg,
                        double lam, int32_t threads, int32_t* status, int32_t* feat, int32_t* bin, double* value,
                        int32_t* count, double* scores) {
  Ctx x;
  if (K < 2 || ctx_init(&x, codes, n, F, NULL, 1, B, D, mcw, msg, lam, NULL, threads)) { ctx_free(&x); return -1; }
  double* g = (double*)malloc((size_t)n * K * sizeof(double));
  double* h = (double*)malloc((size_t)n * K * sizeof(double));
  int32_t* leaf = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  double* cnt = (double*)calloc((size_t)K, sizeof(double));
  if (!g || !h || !leaf || !cnt) { free(g); free(h); free(leaf); free(cnt); ctx_free(&x); return -1; }
  for (int64_t i = 0; i < n; ++i) cnt[labels[i]] += 1.0;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < K; ++k) scores[i * K + k] = log(cnt[k] / (double)n + 1e-10);
  const int NH = heap_nodes(D);
  int rc = 0;
  x.leaf_of_row = leaf;
  for (int t = 0; t < n_rounds && rc == 0; ++t) {
#pragma omp parallel for num_threads(x.threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double* s = scores + i * K;
      double mx = s[0], den = 0.0;
      for (int k = 1; k < K; ++k) mx = s[k] > mx ? s[k] : mx;
      for (int k = 0; k < K; ++k) den += exp(s[k] - mx);
      for (int k = 0; k < K; ++k) {
        const double p = exp(s[k] - mx) / den;
        g[(int64_t)k * n + i] = -(p - (labels[i] == k ? 1.0 : 0.0)) * w[i];
        const double hv = p * (1.0 - p);
        h[(int64_t)k * n + i] = (hv > 1e-6 ? hv : 1e-6) * w[i];
      }
    }
    for (int k = 0; k < K && rc == 0; ++k) {
      const int64_t o = ((int64_t)t * K + k) * NH;
      x.g = g + (int64_t)k * n; x.h = h + (int64_t)k * n;
      x.status = status + o; x.feat = feat + o; x.bin = bin + o; x.value = value + o; x.count = count + o;
      rc = grow_one(&x);
      for (int64_t i = 0; i < n; ++i) scores[i * K + k] += lr * x.value[leaf[i]];
    }
  }
  free(g); free(h); free(leaf); free(cnt);
  ctx_free(&x);
  return rc;
}
