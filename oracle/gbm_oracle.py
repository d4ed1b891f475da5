"""FP64 NumPy restatement of the WarpGBM hot path (test infrastructure only).

Every function cites the reference file:line whose semantics it restates.
Edges and codes are float32 / int8 bit-exact definitions; histograms, gains,
leaf values and predictions are float64.  See ``oracle/__init__.py`` for what
pins each function.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------- binning
def linspace_f32(num_bins: int) -> np.ndarray:
    """Interior quantile levels, ``torch.linspace(0, 1, B+1)[1:-1]`` in float32.

    Follows warpgbm/core.py:1273-1275.  torch's linspace forms
    ``step = f32(1/(B))`` and each level as one fused multiply-add
    ``start + step*i`` (first half) or ``end - step*(steps-1-i)`` (second
    half), rounded once to float32.
    """
    steps = num_bins + 1
    step = F32(F32(1.0) / F32(steps - 1))
    half = steps // 2
    out = np.empty(steps, dtype=F32)
    for i in range(steps):
        if i < half:
            out[i] = F32(0.0 + np.float64(step) * i)
        else:
            out[i] = F32(1.0 - np.float64(step) * (steps - i - 1))
    return out[1:-1]


def sort_total_order(col: np.ndarray) -> np.ndarray:
    """Ascending float32 sort with -0.0 ordered before +0.0 (the order of a
    radix sort on the float bits).  Plain comparison sorts leave the two
    zeros in arbitrary relative order, which can flip the sign of a zero
    edge (never its value, so never a code)."""
    bits = col.view(np.int32)
    key = np.where(bits < 0, ~bits ^ np.int32(-0x80000000), bits)
    return col[np.argsort(key, kind="stable")]


def fma_f32(a, b, c) -> np.float32:
    """Correctly rounded float32 fused multiply-add a*b + c (exact rational
    arithmetic, then one round-to-nearest-even to float32)."""
    from fractions import Fraction
    a, b, c = F32(a), F32(b), F32(c)
    if not (np.isfinite(a) and np.isfinite(b) and np.isfinite(c)):
        return F32(np.float64(a) * np.float64(b) + np.float64(c))
    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    y = F32(float(x))  # within one float32 ulp of x
    best = None
    for cand in (np.nextafter(y, F32(-np.inf)), y, np.nextafter(y, F32(np.inf))):
        err = abs(Fraction(float(cand)) - x)
        if best is None or err < best[0] or (err == best[0] and (int(cand.view(np.int32)) & 1) == 0):
            best = (err, cand)
    r = F32(best[1])
    if r == 0:
        if x == 0:  # exact zero: -0 only when a*b and c are both negative zeros
            neg = (bool(np.signbit(a)) != bool(np.signbit(b))) and bool(np.signbit(c))
        else:       # underflow to zero keeps the sign of the exact result
            neg = x < 0
        r = F32(-0.0) if neg else F32(0.0)
    return r


def quantile_edges_f32(col: np.ndarray, num_bins: int) -> np.ndarray:
    """Bin edges of one feature column: ``torch.quantile(col, q)`` in float32.

    Follows warpgbm/core.py:1271-1278 (linear interpolation).  Definition:
    sort the column (sort_total_order); ``rank = f32(q*(N-1))``;
    ``lo = floor(rank)``, ``hi = ceil(rank)``; ``w = f32(rank - lo)``;
    lerp with ONE rounding per branch, as torch's lerp evaluates it:
    ``w < 0.5 ? fma(w, b-a, a) : fma(-(b-a), 1-w, b)``.  A column holding a
    NaN yields NaN edges (torch propagates NaN through quantile).
    """
    col = np.asarray(col, dtype=F32)
    q = linspace_f32(num_bins)
    n = col.shape[0]
    if n == 0:
        raise ValueError("quantile of an empty column")
    if np.isnan(col).any():
        return np.full(num_bins - 1, np.nan, dtype=F32)
    s = sort_total_order(col)
    out = np.empty(num_bins - 1, dtype=F32)
    for j, qq in enumerate(q):
        rank = F32(qq * F32(n - 1))
        lo = int(np.floor(rank))
        hi = int(np.ceil(rank))
        w = F32(rank - F32(lo))
        a, b = s[lo], s[hi]
        d = F32(b - a)
        if w < F32(0.5):
            out[j] = fma_f32(w, d, a)
        else:
            out[j] = fma_f32(-d, F32(F32(1.0) - w), b)
    return out


def bin_codes(col: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """int8 code of each value: count of leading edges with ``x >= edge``.

    Follows warpgbm/cuda/binner.cu:16-25 (linear scan that stops at the first
    edge greater than x; NaN compares false and yields 0).
    """
    col = np.asarray(col, dtype=F32)
    edges = np.asarray(edges, dtype=F32)
    code = np.zeros(col.shape[0], dtype=np.int64)
    alive = np.ones(col.shape[0], dtype=bool)
    for e in edges:
        step = alive & (col >= e)
        code += step
        alive = step
    return code.astype(np.int8)


def is_prebinned(X: np.ndarray, num_bins: int) -> bool:
    """Pre-binned bypass test, warpgbm/core.py:1242-1245."""
    return bool(np.issubdtype(X.dtype, np.integer) and np.all(X.max(axis=0) < num_bins))


def bin_matrix(X: np.ndarray, num_bins: int):
    """Codes [N,F] int8 and edges [F,B-1] f32, warpgbm/core.py:1226-1285."""
    X = np.asarray(X)
    N, Fn = X.shape
    if is_prebinned(X, num_bins):
        edges = np.tile(np.arange(1, num_bins, dtype=F32), (Fn, 1))
        return X.astype(np.int8), edges
    Xf = X.astype(F32)
    codes = np.empty((N, Fn), dtype=np.int8)
    edges = np.empty((Fn, num_bins - 1), dtype=F32)
    for f in range(Fn):
        edges[f] = quantile_edges_f32(Xf[:, f], num_bins)
        codes[:, f] = bin_codes(Xf[:, f], edges[f])
    return codes, edges


def bin_with_edges(X: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Inference-time binning with stored edges, warpgbm/core.py:1882-1900."""
    X = np.asarray(X, dtype=F32)
    out = np.empty(X.shape, dtype=np.int8)
    for f in range(X.shape[1]):
        out[:, f] = bin_codes(X[:, f], edges[f])
    return out


def era_inverse(era_id: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """``torch.unique(era, return_inverse=True)``, warpgbm/core.py:1284."""
    uniq, inv = np.unique(np.asarray(era_id).astype(np.int32), return_inverse=True)
    return uniq, inv.astype(np.int32)


# ---------------------------------------------------------------- histogram
def histogram(codes, g, h, rows, feats, eras, num_eras, num_bins):
    """G, H [E, F_sel, B] float64 over the listed rows.

    Follows warpgbm/cuda/histogram_kernel.cu:33-48: for each selected
    feature j (global column feats[j]) and row r, if 0 <= code < B then
    G[era[r], j, code] += g[r] (same for H).  Codes outside [0, B) drop.
    """
    rows = np.asarray(rows, dtype=np.int64)
    feats = np.asarray(feats, dtype=np.int64)
    E, B = int(num_eras), int(num_bins)
    G = np.zeros((E, len(feats), B), dtype=np.float64)
    H = np.zeros((E, len(feats), B), dtype=np.float64)
    if rows.size == 0:
        return G, H
    gr = np.asarray(g, dtype=np.float64)[rows]
    hr = np.asarray(h, dtype=np.float64)[rows]
    er = np.asarray(eras, dtype=np.int64)[rows]
    sub = codes[rows][:, feats].astype(np.int64)
    for j in range(len(feats)):
        c = sub[:, j]
        ok = (c >= 0) & (c < B)
        key = er[ok] * B + c[ok]
        G[:, j, :] = np.bincount(key, weights=gr[ok], minlength=E * B).reshape(E, B)
        H[:, j, :] = np.bincount(key, weights=hr[ok], minlength=E * B).reshape(E, B)
    return G, H


# ---------------------------------------------------------------- split scan
def split_gains(G, H, min_child_weight, eps):
    """Per-era gain and direction [E, F, B-1] (float64).

    Follows warpgbm/cuda/best_split_kernel.cu:25-54: prefix G_L, H_L over
    bins b < B-1; if H_L >= mcw and H_R >= mcw then
    gain = G_L^2/(H_L+eps) + G_R^2/(H_R+eps) - G^2/(H+eps) and
    dir = +1 if G_L/(H_L+eps) > G_R/(H_R+eps) else -1; otherwise 0, 0.
    """
    G = np.asarray(G, dtype=np.float64)
    H = np.asarray(H, dtype=np.float64)
    Gt = G.sum(axis=-1, keepdims=True)
    Ht = H.sum(axis=-1, keepdims=True)
    GL = np.cumsum(G, axis=-1)[..., :-1]
    HL = np.cumsum(H, axis=-1)[..., :-1]
    GR = Gt - GL
    HR = Ht - HL
    ok = (HL >= min_child_weight) & (HR >= min_child_weight)
    with np.errstate(divide="ignore", invalid="ignore"):
        gain = GL * GL / (HL + eps) + GR * GR / (HR + eps) - Gt * Gt / (Ht + eps)
        lv = GL / (HL + eps)
        rv = GR / (HR + eps)
    d = np.where(lv > rv, 1.0, -1.0)
    return np.where(ok, gain, 0.0), np.where(ok, d, 0.0)


def select_split(gain, direction, min_split_gain, mono_required_local=None):
    """DES selection of one (local feature, bin), or (-1, -1) for a leaf.

    Follows warpgbm/core.py:1322-1362:
      E == 1: crit = gain[0], mask = crit > msg.
      E  > 1: agree = |mean_e dir|, crit = mean_e gain,
              mask = (agree == max agree over ALL (f,b)) & (crit > msg).
      Monotone (required dir r != 0 for the feature): E == 1 needs dir == r;
      E > 1 needs dir == r in every era.
      No candidate -> (-1, -1); else the first index of the max crit over the
      flattened [F, B-1] grid.
    The agreement test uses the integer |sum_e dir| (same order as the mean).
    """
    gain = np.asarray(gain, dtype=np.float64)
    direction = np.asarray(direction, dtype=np.float64)
    E, Fl, Bm1 = gain.shape
    if E == 1:
        crit = gain[0].copy()
        mask = crit > min_split_gain
    else:
        sdir = np.abs(np.rint(direction.sum(axis=0))).astype(np.int64)
        crit = gain.mean(axis=0)
        mask = (sdir == sdir.max()) & (crit > min_split_gain)
    if mono_required_local is not None:
        req = np.asarray(mono_required_local, dtype=np.float64).reshape(-1, 1)
        row = req != 0
        if row.any():
            if E == 1:
                mm = (~row) | (direction[0] == req)
            else:
                mm = (~row) | np.all(direction == req[None, :, :], axis=0)
            mask = mask & mm
    if not mask.any():
        return -1, -1
    crit = np.where(mask, crit, -np.inf)
    idx = int(np.argmax(crit))
    return idx // Bm1, idx % Bm1


def leaf_value(g, h, rows, lam):
    """sum(g)/(sum(h)+lambda) over the node's rows, warpgbm/core.py:1377-1380."""
    rows = np.asarray(rows, dtype=np.int64)
    return float(np.sum(np.asarray(g, np.float64)[rows]) /
                 (np.sum(np.asarray(h, np.float64)[rows]) + lam))


# ---------------------------------------------------------------- tree growth
class TreeRecord:
    """Result of growing one tree: nested dict plus flat bookkeeping.

    ``nodes`` maps heap id (root 0, children 2i+1 / 2i+2) to
    ("split", feature, bin, per_era_gain) or ("leaf", value, samples).
    ``leaf_of_row`` holds the heap id of the leaf each listed row ends in.
    """

    def __init__(self):
        self.tree = None
        self.nodes = {}
        self.leaf_of_row = None


def grow_tree(codes, g, h, eras, num_eras, num_bins, max_depth, min_child_weight,
              min_split_gain, lam, feats=None, mono_required=None, rows=None):
    """Depth-first tree growth, warpgbm/core.py:1364-1470.

    ``g``/``h`` are the (weighted) gradient and hessian the tree is fitted
    to; the caller may pass the exact float32 values a device build used
    (teacher forcing).  Children are ordered left (code <= bin) then right,
    each keeping the parent's row order (boolean-mask compaction,
    core.py:1425-1427).  Returns a TreeRecord.
    """
    N, Fm = codes.shape
    feats = np.arange(Fm) if feats is None else np.asarray(feats, dtype=np.int64)
    rows = np.arange(N) if rows is None else np.asarray(rows, dtype=np.int64)
    mono_local = None
    if mono_required is not None:
        mono_local = np.asarray(mono_required, dtype=np.int64)[feats]
    rec = TreeRecord()
    rec.leaf_of_row = np.full(N, -1, dtype=np.int64)

    def make_leaf(node_rows, heap):
        v = leaf_value(g, h, node_rows, lam)
        rec.nodes[heap] = ("leaf", v, int(node_rows.size))
        rec.leaf_of_row[node_rows] = heap
        return {"leaf_value": v, "samples": int(node_rows.size)}

    def grow(node_rows, depth, heap):
        if depth == max_depth:
            return make_leaf(node_rows, heap)
        G, H = histogram(codes, g, h, node_rows, feats, eras, num_eras, num_bins)
        gain, d = split_gains(G, H, min_child_weight, lam)
        fl, b = select_split(gain, d, min_split_gain, mono_local)
        if fl == -1:
            return make_leaf(node_rows, heap)
        feat = int(feats[fl])
        rec.nodes[heap] = ("split", feat, int(b), gain[:, fl, b].copy())
        go_left = codes[node_rows, feat] <= b
        left = grow(node_rows[go_left], depth + 1, 2 * heap + 1)
        right = grow(node_rows[~go_left], depth + 1, 2 * heap + 2)
        return {"feature": feat, "bin": int(b), "left": left, "right": right}

    rec.tree = grow(rows, 0, 0)
    return rec


# ---------------------------------------------------------------- objectives
def softmax(logits):
    """Max-subtracted softmax, warpgbm/metrics.py:12-15 (float64)."""
    z = np.asarray(logits, np.float64)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def softmax_grad_hess(scores, y, w):
    """Per-class tree targets for one boosting round.

    Follows warpgbm/core.py:339-367 and :1755-1756:
    p = softmax(F); grad_k = -(p_k - onehot_k) * w;
    hess_k = max(p_k (1 - p_k), 1e-6) * w.  Returns [N, K] each.
    """
    p = softmax(scores)
    K = p.shape[1]
    onehot = np.eye(K)[np.asarray(y, dtype=np.int64)]
    w = np.asarray(w, np.float64)[:, None]
    return -(p - onehot) * w, np.maximum(p * (1.0 - p), 1e-6) * w


def fit_regression(codes, y, w, eras, num_eras, num_bins, max_depth, n_trees, lr,
                   min_child_weight=20, min_split_gain=0.0, lam=1e-6, mono_required=None):
    """Boosting loop for the regression objective, warpgbm/core.py:1518-1580.

    base = unweighted mean of y (core.py:926); per tree g = (y - F) w,
    h = w (core.py:1545-1547); leaves add lr*value to F.
    Returns (base, list of TreeRecord, final scores F).
    """
    y = np.asarray(y, np.float64)
    w = np.asarray(w, np.float64)
    base = float(np.mean(y))
    scores = np.full(y.shape[0], base)
    recs = []
    for _ in range(n_trees):
        g = (y - scores) * w
        h = w.copy()
        rec = grow_tree(codes, g, h, eras, num_eras, num_bins, max_depth,
                        min_child_weight, min_split_gain, lam, mono_required=mono_required)
        vals = np.array([rec.nodes[int(k)][1] for k in rec.leaf_of_row])
        scores = scores + lr * vals
        recs.append(rec)
    return base, recs, scores


def multi_regression_targets(Y, scores, W):
    """Per-output tree targets of the multi-output objective (float64).

    Follows warpgbm/core.py:1666-1673: a target is valid where it is finite;
    g = (y - F) * w and h = w on valid entries, 0 elsewhere.  Returns
    (g, h), both [N, M].
    """
    Y = np.asarray(Y, np.float64)
    valid = np.isfinite(Y)
    w = np.where(valid, np.asarray(W, np.float64), 0.0)
    g = np.where(valid, Y - np.asarray(scores, np.float64), 0.0) * w
    return g, w


def grow_tree_multi(codes, g, h, eras, num_eras, num_bins, max_depth, min_child_weight, min_split_gain, lam,
                    feats=None, mono_required=None, split_g=None, split_h=None):
    """Shared-split tree with a vector of leaf values (float64).

    Follows warpgbm/core.py:1674-1679 (splits searched on the per-row sums
    over outputs of g and h) and :1366-1375 (leaf m = sum g_m / (sum h_m +
    lambda) over the leaf's rows).  ``split_g``/``split_h`` replace the
    row sums when given (teacher forcing with a device build's float32
    sums).  Returns a TreeRecord whose leaves hold lists of M values.
    """
    g = np.asarray(g, np.float64)
    h = np.asarray(h, np.float64)
    sg = g.sum(axis=1) if split_g is None else split_g
    sh = h.sum(axis=1) if split_h is None else split_h
    rec = grow_tree(codes, sg, sh, eras, num_eras, num_bins, max_depth, min_child_weight, min_split_gain, lam,
                    feats=feats, mono_required=mono_required)
    M = g.shape[1]

    def vec(heap):
        rows = np.nonzero(rec.leaf_of_row == heap)[0]
        return [leaf_value(g[:, m], h[:, m], rows, lam) for m in range(M)]

    def walk(node, heap):
        if "leaf_value" in node:
            v = vec(heap)
            node["leaf_value"] = v
            rec.nodes[heap] = ("leaf", v, rec.nodes[heap][2])
            return
        walk(node["left"], 2 * heap + 1)
        walk(node["right"], 2 * heap + 2)

    walk(rec.tree, 0)
    return rec


def fit_multi_regression(codes, Y, W, eras, num_eras, num_bins, max_depth, n_trees, lr, min_child_weight=20,
                         min_split_gain=0.0, lam=1e-6):
    """Boosting loop of the multi-output objective, warpgbm/core.py:1619-1717.

    base_m = weighted mean of output m's valid targets, 0 when it has no
    weight (core.py:1055-1063); each round grows one shared-split tree and
    adds lr * leaf vector to every row's scores.  Returns (base [M], list
    of TreeRecord, final scores [N, M]).
    """
    Y = np.asarray(Y, np.float64)
    valid = np.isfinite(Y)
    w = np.where(valid, np.asarray(W, np.float64), 0.0)
    num = (np.where(valid, Y, 0.0) * w).sum(axis=0)
    den = w.sum(axis=0)
    base = np.where(den > 0, num / np.maximum(den, 1e-12), 0.0)
    scores = np.tile(base, (Y.shape[0], 1))
    recs = []
    for _ in range(n_trees):
        g, h = multi_regression_targets(Y, scores, W)
        rec = grow_tree_multi(codes, g, h, eras, num_eras, num_bins, max_depth, min_child_weight, min_split_gain,
                              lam)
        vals = np.array([rec.nodes[int(k)][1] for k in rec.leaf_of_row], dtype=np.float64)
        scores = scores + lr * vals
        recs.append(rec)
    return base, recs, scores


def per_era_importance(records, num_eras, num_features):
    """Per-era gain importance [E, F] of grown trees (float64).

    Follows warpgbm/core.py:1418-1423: every split adds the per-era gains
    of its chosen (feature, bin) to that feature's column, for every era.
    ``records``: TreeRecords from grow_tree (their split nodes carry the
    per-era gain vector of the winning threshold).
    """
    imp = np.zeros((int(num_eras), int(num_features)), dtype=np.float64)
    for rec in records:
        for node in rec.nodes.values():
            if node[0] == "split":
                imp[:, node[1]] += np.asarray(node[3], dtype=np.float64)
    return imp


# ---------------------------------------------------------------- prediction
def flatten_tree(tree, max_nodes):
    """Preorder flattening to [max_nodes, 6], warpgbm/core.py:2283-2328.

    Row = [feature, bin, left_id, right_id, is_leaf, value]; unused rows NaN.
    """
    flat = np.full((max_nodes, 6), np.nan, dtype=F32)
    counter = [0]

    def walk(node):
        cur = counter[0]
        counter[0] += 1
        if "leaf_value" in node:
            flat[cur, 4] = 1.0
            flat[cur, 5] = float(node["leaf_value"])
        else:
            flat[cur, 0] = float(node["feature"])
            flat[cur, 1] = float(node["bin"])
            flat[cur, 4] = 0.0
            flat[cur, 2] = counter[0]
            walk(node["left"])
            flat[cur, 3] = counter[0]
            walk(node["right"])

    walk(tree)
    return flat


def predict_tree(tree, codes):
    """Leaf value per row of a nested-dict tree (float64)."""
    out = np.empty(codes.shape[0], dtype=np.float64)

    def walk(node, rows):
        if rows.size == 0:
            return
        if "leaf_value" in node:
            out[rows] = node["leaf_value"]
            return
        m = codes[rows, int(node["feature"])] <= int(node["bin"])
        walk(node["left"], rows[m])
        walk(node["right"], rows[~m])

    walk(tree, np.arange(codes.shape[0]))
    return out


def predict_multi(codes, trees, lr, base):
    """[N, M] = base + sum over trees of lr * leaf vector (float64), the
    multi-output prediction of warpgbm/core.py:1904-1935 and the score
    update of :1366-1375."""
    base = np.asarray(base, np.float64).reshape(-1)
    out = np.tile(base, (codes.shape[0], 1))
    for tree in trees:
        def walk(node, rows):
            if rows.size == 0:
                return
            if "leaf_value" in node:
                out[rows] += float(lr) * np.asarray(node["leaf_value"], np.float64)
                return
            m = codes[rows, int(node["feature"])] <= int(node["bin"])
            walk(node["left"], rows[m])
            walk(node["right"], rows[~m])

        walk(tree, np.arange(codes.shape[0]))
    return out


def predict_forest_flat(codes, flat_trees, lr, base=0.0):
    """base + sum_t lr * leaf_t(row) over flattened trees (float64).

    Follows warpgbm/cuda/predict.cu:18-44: from node 0, go left iff
    code[row, feature] <= bin, until is_leaf > 0.5.
    """
    N = codes.shape[0]
    out = np.full(N, float(base), dtype=np.float64)
    rows = np.arange(N)
    for flat in flat_trees:
        node = np.zeros(N, dtype=np.int64)
        while True:
            leaf = flat[node, 4] > 0.5
            if leaf.all():
                break
            nl = ~leaf
            f = flat[node[nl], 0].astype(np.int64)
            b = flat[node[nl], 1].astype(np.int64)
            c = codes[rows[nl], f].astype(np.int64)
            node[nl] = np.where(c <= b, flat[node[nl], 2], flat[node[nl], 3]).astype(np.int64)
        out += lr * flat[node, 5].astype(np.float64)
    return out


def predict_proba_binned(codes, forest_rounds, lr, num_classes):
    """Class probabilities of a multiclass forest on bin codes [N, K].

    Follows warpgbm/core.py:2021-2058: raw score F[:, k] starts at the
    uniform prior log(1/K + 1e-10) (computed in float64, stored in a float32
    tensor), every class-k tree adds lr * leaf, and the probabilities are the
    max-subtracted softmax of F (metrics.py:12-15).  ``forest_rounds``: a
    list of boosting rounds, each a list of K nested-dict trees (the
    estimator's ``forest`` layout for classification).
    """
    K = int(num_classes)
    N = codes.shape[0]
    prior = float(np.float32(np.log(1.0 / K + 1e-10)))
    F = np.full((N, K), prior, dtype=np.float64)
    for trees in forest_rounds:
        for k in range(K):
            F[:, k] += float(lr) * predict_tree(trees[k], codes)
    return softmax(F)
