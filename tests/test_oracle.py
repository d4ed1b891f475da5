"""Pins the FP64 oracle (CPU only).

The reference ships no golden vectors for this path (its tests are
end-to-end), so each oracle function is checked against an independent
definition: torch's own CPU float32 linspace/quantile (the library routine
the reference calls), np.searchsorted, brute-force loops on tiny inputs,
closed forms and invariants.  test_planted_mistakes_are_caught shows that
plausible slips in the oracle fail these checks.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import gbm_oracle as O


def test_linspace_matches_torch_for_every_bin_count():
    for B in range(2, 128):
        ref = torch.linspace(0, 1, B + 1, dtype=torch.float32).numpy()[1:-1]
        assert np.array_equal(O.linspace_f32(B), ref), B


@pytest.mark.parametrize("N", [2, 7, 1000, 12345, 20000])
@pytest.mark.parametrize("B", [2, 5, 10, 32, 127])
def test_quantile_edges_match_torch_quantile(N, B):
    rng = np.random.default_rng(N * 131 + B)
    for col in (rng.standard_normal(N).astype(np.float32),
                np.round(rng.standard_normal(N) * 2).astype(np.float32),   # ties
                (rng.standard_normal(N) * 1e-30).astype(np.float32)):
        q = torch.linspace(0, 1, B + 1, dtype=torch.float32)[1:-1]
        ref = torch.quantile(torch.from_numpy(col), q).numpy()
        got = O.quantile_edges_f32(col, B)
        # equal values; equal bits except possibly the sign of a zero edge,
        # which torch's comparison sort leaves to chance (codes unaffected)
        assert np.array_equal(got, ref)
        nz = got != 0
        assert np.array_equal(got[nz].view(np.int32), ref[nz].view(np.int32))


def test_quantile_edges_nan_column():
    col = np.array([1.0, np.nan, 3.0, 2.0], np.float32)
    q = torch.linspace(0, 1, 5, dtype=torch.float32)[1:-1]
    ref = torch.quantile(torch.from_numpy(col), q).numpy()
    assert np.all(np.isnan(ref)) and np.all(np.isnan(O.quantile_edges_f32(col, 4)))


def test_quantile_edges_large_n_float32_rank():
    """At N = 1M the float32 rank q*(N-1) rounds (e.g. 968749.03 -> 968749):
    the oracle must follow torch's float32 arithmetic, not float64."""
    rng = np.random.default_rng(0)
    col = rng.standard_normal(1_000_000).astype(np.float32)
    q = torch.linspace(0, 1, 33, dtype=torch.float32)[1:-1]
    ref = torch.quantile(torch.from_numpy(col), q).numpy()
    assert np.array_equal(O.quantile_edges_f32(col, 32), ref)
    f64 = np.quantile(col.astype(np.float64), q.numpy().astype(np.float64)).astype(np.float32)
    assert not np.array_equal(f64, ref)  # the naive float64 definition differs here


def test_bin_codes_match_searchsorted_and_nan():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(10000).astype(np.float32)
    x[::97] = np.nan
    edges = np.sort(rng.standard_normal(31)).astype(np.float32)
    edges[5] = edges[6]  # duplicate edge
    codes = O.bin_codes(x, edges)
    ok = ~np.isnan(x)
    assert np.array_equal(codes[ok], np.searchsorted(edges, x[ok], side="right"))
    assert np.all(codes[~ok] == 0)


def test_prebinned_detection():
    assert O.is_prebinned(np.array([[0, 1], [4, 2]], np.int64), 5)
    assert not O.is_prebinned(np.array([[0, 1], [5, 2]], np.int64), 5)
    assert not O.is_prebinned(np.array([[0.0, 1.0]], np.float32), 5)


def _brute_hist(codes, g, h, rows, feats, eras, E, B):
    G = np.zeros((E, len(feats), B))
    H = np.zeros((E, len(feats), B))
    for r in rows:
        for j, f in enumerate(feats):
            c = int(codes[r, f])
            if 0 <= c < B:
                G[eras[r], j, c] += g[r]
                H[eras[r], j, c] += h[r]
    return G, H


def test_histogram_brute_force_and_invariants():
    rng = np.random.default_rng(2)
    N, F, B, E = 300, 6, 5, 3
    codes = rng.integers(-1, B + 1, size=(N, F)).astype(np.int8)
    g = rng.standard_normal(N)
    h = rng.uniform(0.5, 2, N)
    eras = rng.integers(0, E, N)
    rows = rng.choice(N, 200, replace=False)
    feats = np.array([4, 1, 3])
    G, H = O.histogram(codes, g, h, rows, feats, eras, E, B)
    Gb, Hb = _brute_hist(codes, g, h, rows, feats, eras, E, B)
    np.testing.assert_allclose(G, Gb, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(H, Hb, rtol=1e-12, atol=1e-12)
    # parent = left + right
    m = codes[rows, 2] <= 2
    Gl, _ = O.histogram(codes, g, h, rows[m], feats, eras, E, B)
    Gr, _ = O.histogram(codes, g, h, rows[~m], feats, eras, E, B)
    np.testing.assert_allclose(G, Gl + Gr, atol=1e-12)


def _brute_gain(codes, g, h, rows, f, b, mcw, lam):
    """Gain of threshold b on feature f from the row subsets directly."""
    c = codes[rows, f].astype(int)
    ok = (c >= 0)
    left = rows[ok & (c <= b)]
    right = rows[ok & (c > b)]
    allr = rows[ok]
    GL, HL = g[left].sum(), h[left].sum()
    GR, HR = g[right].sum(), h[right].sum()
    if HL < mcw or HR < mcw:
        return 0.0, 0.0
    G, H = g[allr].sum(), h[allr].sum()
    gain = GL * GL / (HL + lam) + GR * GR / (HR + lam) - G * G / (H + lam)
    return gain, (1.0 if GL / (HL + lam) > GR / (HR + lam) else -1.0)


def test_split_gains_brute_force():
    rng = np.random.default_rng(3)
    N, F, B = 400, 4, 6
    codes = rng.integers(0, B, size=(N, F)).astype(np.int8)
    g = rng.standard_normal(N) + (codes[:, 1] > 2)
    h = rng.uniform(0.5, 2, N)
    rows = np.arange(N)
    G, H = O.histogram(codes, g, h, rows, np.arange(F), np.zeros(N, int), 1, B)
    gain, d = O.split_gains(G, H, 20.0, 1e-6)
    for f, b in itertools.product(range(F), range(B - 1)):
        bg, bd = _brute_gain(codes, g, h, rows, f, b, 20.0, 1e-6)
        assert abs(gain[0, f, b] - bg) < 1e-9 * max(1, abs(bg))
        assert d[0, f, b] == bd


def _brute_select(gain, d, msg, req=None):
    """Reference selection written as a plain loop over every (f, b)."""
    E, F, Bm1 = gain.shape
    cands = []
    for f in range(F):
        for b in range(Bm1):
            if E == 1:
                crit, agree = gain[0, f, b], 0
            else:
                crit = sum(gain[e, f, b] for e in range(E)) / E
                agree = abs(sum(d[e, f, b] for e in range(E)))
            mono = True
            if req is not None and req[f] != 0:
                mono = all(d[e, f, b] == req[f] for e in range(E))
            cands.append((agree, crit, f, b, mono))
    top = max(c[0] for c in cands)
    best = None
    for agree, crit, f, b, mono in cands:
        if agree == top and crit > msg and mono and (best is None or crit > best[0]):
            best = (crit, f, b)
    return (-1, -1) if best is None else (best[1], best[2])


def test_select_split_brute_force():
    rng = np.random.default_rng(4)
    for trial in range(200):
        E = int(rng.integers(1, 5))
        F, Bm1 = int(rng.integers(1, 5)), int(rng.integers(1, 6))
        gain = np.round(rng.standard_normal((E, F, Bm1)), 1)  # rounding creates ties
        d = rng.choice([-1.0, 0.0, 1.0], size=(E, F, Bm1))
        req = rng.choice([-1, 0, 1], size=F) if trial % 2 else None
        msg = float(rng.choice([0.0, 0.3]))
        assert O.select_split(gain, d, msg, req) == _brute_select(gain, d, msg, req), trial


def test_select_split_special_cases():
    # conflicting eras under a monotone constraint -> leaf
    gain = np.ones((2, 1, 4))
    d = np.stack([np.ones((1, 4)), -np.ones((1, 4))])
    assert O.select_split(gain, d, 0.0, np.array([-1])) == (-1, -1)
    # a higher agreement wins even when it has no positive gain -> leaf
    gain = np.array([[[5.0, -1.0]], [[5.0, -1.0]]])
    d = np.array([[[1.0, 1.0]], [[-1.0, 1.0]]])
    assert O.select_split(gain, d, 0.0) == (-1, -1)
    # mcw larger than the node -> every gain 0 -> leaf
    G = np.ones((1, 3, 4))
    gain0, d0 = O.split_gains(G, G, 100.0, 1e-6)
    assert O.select_split(gain0, d0, 0.0) == (-1, -1)


def test_grow_tree_assignment_and_leaves():
    rng = np.random.default_rng(5)
    N, F, B = 3000, 5, 8
    codes = rng.integers(0, B, size=(N, F)).astype(np.int8)
    g = rng.standard_normal(N) + 2 * (codes[:, 0] > 3)
    h = np.ones(N)
    rec = O.grow_tree(codes, g, h, np.zeros(N, int), 1, B, 3, 20, 0.0, 1e-6)
    vals = O.predict_tree(rec.tree, codes)
    for heap, node in rec.nodes.items():
        if node[0] == "leaf":
            rows = np.where(rec.leaf_of_row == heap)[0]
            assert rows.size == node[2]
            assert abs(node[1] - g[rows].sum() / (h[rows].sum() + 1e-6)) < 1e-12
            assert np.all(vals[rows] == node[1])
    assert np.all(rec.leaf_of_row >= 0)
    # the root split is the brute-force best split
    root = rec.nodes[0]
    assert root[0] == "split"
    best = max(((f, b) for f in range(F) for b in range(B - 1)),
               key=lambda fb: (_brute_gain(codes, g, h, np.arange(N), fb[0], fb[1], 20, 1e-6)[0], -fb[0], -fb[1]))
    assert (root[1], root[2]) == best


def test_flatten_and_flat_predict_match_dict_traversal():
    rng = np.random.default_rng(6)
    N, F, B = 2000, 6, 8
    codes = rng.integers(0, B, size=(N, F)).astype(np.int8)
    trees = [O.grow_tree(codes, rng.standard_normal(N) + (codes[:, k] > 4), np.ones(N), np.zeros(N, int), 1, B,
                         4, 20, 0.0, 1e-6).tree for k in range(4)]
    flats = [O.flatten_tree(t, 32) for t in trees]
    ref = 0.25 + 0.1 * sum(O.predict_tree(t, codes) for t in trees)
    np.testing.assert_allclose(O.predict_forest_flat(codes, flats, 0.1, base=0.25), ref, rtol=1e-6, atol=1e-7)
    half = O.predict_forest_flat(codes[:1000], flats, 0.1, base=0.25)
    assert np.array_equal(half, O.predict_forest_flat(codes, flats, 0.1, base=0.25)[:1000])


def test_softmax_grad_hess_closed_forms():
    rng = np.random.default_rng(7)
    s = rng.standard_normal((100, 2))
    y = rng.integers(0, 2, 100)
    w = rng.uniform(0.5, 1.5, 100)
    g, h = O.softmax_grad_hess(s, y, w)
    p1 = 1.0 / (1.0 + np.exp(-(s[:, 1] - s[:, 0])))
    np.testing.assert_allclose(g[:, 1], -(p1 - (y == 1)) * w, rtol=1e-12)
    np.testing.assert_allclose(h[:, 1], np.maximum(p1 * (1 - p1), 1e-6) * w, rtol=1e-12)
    np.testing.assert_allclose(O.softmax(rng.standard_normal((50, 5))).sum(axis=1), 1.0, rtol=1e-12)
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-12)


def test_planted_mistakes_are_caught(monkeypatch):
    """A dropped parent term in the gain, or a flipped direction, fails the
    brute-force pins above."""
    rng = np.random.default_rng(3)
    N, F, B = 400, 4, 6
    codes = rng.integers(0, B, size=(N, F)).astype(np.int8)
    g = rng.standard_normal(N) + (codes[:, 1] > 2)
    h = rng.uniform(0.5, 2, N)
    G, H = O.histogram(codes, g, h, np.arange(N), np.arange(F), np.zeros(N, int), 1, B)

    def dropped_parent(G, H, mcw, eps):
        GL = np.cumsum(G, -1)[..., :-1]
        HL = np.cumsum(H, -1)[..., :-1]
        GR = G.sum(-1, keepdims=True) - GL
        HR = H.sum(-1, keepdims=True) - HL
        return GL * GL / (HL + eps) + GR * GR / (HR + eps), np.where(GL / HL > GR / HR, -1.0, 1.0)

    bad_gain, bad_dir = dropped_parent(G, H, 20.0, 1e-6)
    bg, bd = _brute_gain(codes, g, h, np.arange(N), 1, 2, 20.0, 1e-6)
    assert abs(bad_gain[0, 1, 2] - bg) > 1e-3
    assert bad_dir[0, 1, 2] != bd


def test_era_inverse_matches_torch_unique():
    """era_inverse restates torch.unique(era, return_inverse=True)
    (warpgbm/core.py:1284); checked against torch's CPU routine."""
    rng = np.random.default_rng(12)
    for era in (rng.integers(-5, 40, 3000), rng.choice([7, 3, 1000, -2], 500), np.zeros(10, np.int64)):
        uniq, inv = O.era_inverse(era)
        tu, ti = torch.unique(torch.from_numpy(np.asarray(era).astype(np.int32)), return_inverse=True)
        assert np.array_equal(uniq, tu.numpy()) and np.array_equal(inv, ti.numpy())
        assert inv.dtype == np.int32


def test_fit_regression_single_leaf_closed_form():
    """With min_child_weight above the row count every tree is one leaf and
    F_t = ybar_w + (ybar - ybar_w) * (1 - lr * S_w / (S_w + lambda))^t, with
    the UNWEIGHTED base ybar (warpgbm/core.py:926), g = (y - F) w and h = w
    (core.py:1545-1547).  A weighted base, a flipped gradient sign or a
    dropped lambda all change this sequence."""
    rng = np.random.default_rng(13)
    N, F, B = 500, 4, 8
    codes = rng.integers(0, B, (N, F)).astype(np.int8)
    y = rng.standard_normal(N) * 3 + 1.5
    w = rng.uniform(0.2, 3.0, N)
    lr, lam, T = 0.3, 40.0, 7
    base, recs, scores = O.fit_regression(codes, y, w, np.zeros(N, int), 1, B, 3, T, lr,
                                          min_child_weight=w.sum() + 1.0, lam=lam)
    ybar, yw, Sw = y.mean(), (w * y).sum() / w.sum(), w.sum()
    assert base == ybar
    for t, rec in enumerate(recs):
        assert list(rec.nodes) == [0] and rec.nodes[0][0] == "leaf"
        Ft = yw + (ybar - yw) * (1 - lr * Sw / (Sw + lam)) ** t
        assert abs(rec.nodes[0][1] - Sw * (yw - Ft) / (Sw + lam)) < 1e-12 * max(1.0, abs(Ft))
    FT = yw + (ybar - yw) * (1 - lr * Sw / (Sw + lam)) ** T
    np.testing.assert_allclose(scores, FT, rtol=1e-12)


def _node_rows(tree, codes, rows=None, heap=0, out=None):
    """Rows reaching each heap node of a nested-dict tree."""
    if out is None:
        out, rows = {}, np.arange(codes.shape[0])
    out[heap] = rows
    if "leaf_value" not in tree:
        m = codes[rows, tree["feature"]] <= tree["bin"]
        _node_rows(tree["left"], codes, rows[m], 2 * heap + 1, out)
        _node_rows(tree["right"], codes, rows[~m], 2 * heap + 2, out)
    return out


def test_per_era_importance_brute_force():
    """Each split's era-e gain recomputed from the node's era-e rows alone
    (direct sums, no histogram) and added per feature equals
    per_era_importance (warpgbm/core.py:1418-1423); summed over eras of a
    one-era fit it is the total split gain."""
    rng = np.random.default_rng(14)
    N, F, B, E = 3000, 5, 6, 3
    codes = rng.integers(0, B, (N, F)).astype(np.int8)
    eras = rng.integers(0, E, N)
    g = rng.standard_normal(N) + 1.5 * (codes[:, 2] > 2) - (codes[:, 4] > 3)
    h = rng.uniform(0.5, 1.5, N)
    recs = [O.grow_tree(codes, g * s, h, eras, E, B, 3, 20, -np.inf, 1e-6) for s in (1.0, -0.7)]
    imp = O.per_era_importance(recs, E, F)
    ref = np.zeros((E, F))
    n_splits = 0
    for rec in recs:
        nodes = _node_rows(rec.tree, codes)
        for heap, node in rec.nodes.items():
            if node[0] != "split":
                continue
            n_splits += 1
            for e in range(E):
                r = nodes[heap][eras[nodes[heap]] == e]
                ref[e, node[1]] += _brute_gain(codes, g * (1.0 if rec is recs[0] else -0.7), h, r, node[1],
                                               node[2], 20, 1e-6)[0]
    assert n_splits >= 6
    np.testing.assert_allclose(imp, ref, rtol=1e-9, atol=1e-9)
    one = O.grow_tree(codes, g, h, np.zeros(N, int), 1, B, 3, 20, 0.0, 1e-6)
    tot = sum(node[3][0] for node in one.nodes.values() if node[0] == "split")
    assert abs(O.per_era_importance([one], 1, F).sum() - tot) < 1e-9 * abs(tot)


def test_predict_proba_binned_closed_forms():
    """K = 2: p1 = sigmoid(lr * (s1 - s0)) with s_k the class-k forest sums
    (the uniform prior cancels); rows sum to 1; permuting the classes of
    every round permutes the columns."""
    rng = np.random.default_rng(15)
    N, F, B = 2000, 6, 8
    codes = rng.integers(0, B, (N, F)).astype(np.int8)

    def tree(seed):
        r = np.random.default_rng(seed)
        return O.grow_tree(codes, r.standard_normal(N) + 2 * (codes[:, seed % F] > 3), np.ones(N),
                           np.zeros(N, int), 1, B, 3, 20, 0.0, 1e-6).tree

    rounds = [[tree(1), tree(2)], [tree(3), tree(4)]]
    p = O.predict_proba_binned(codes, rounds, 0.1, 2)
    s = [sum(O.predict_tree(rd[k], codes) for rd in rounds) for k in range(2)]
    np.testing.assert_allclose(p[:, 1], 1.0 / (1.0 + np.exp(-0.1 * (s[1] - s[0]))), rtol=1e-12)
    rounds4 = [[tree(10 + 4 * i + k) for k in range(4)] for i in range(3)]
    p4 = O.predict_proba_binned(codes, rounds4, 0.2, 4)
    np.testing.assert_allclose(p4.sum(axis=1), 1.0, rtol=1e-12)
    perm = [2, 0, 3, 1]
    p4p = O.predict_proba_binned(codes, [[rd[j] for j in perm] for rd in rounds4], 0.2, 4)
    np.testing.assert_allclose(p4p, p4[:, perm], rtol=1e-12)


def _multi_data(seed, N=1500, F=6, B=16, M=3):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, B, (N, F)).astype(np.int8)
    Y = np.stack([2.0 * (codes[:, m % F] > B // 2) + codes[:, (m + 1) % F] * 0.3 * (m + 1) + rng.standard_normal(N)
                  for m in range(M)], axis=1)
    return rng, codes, Y


def _same_structure(a, b):
    if ("leaf_value" in a) != ("leaf_value" in b):
        return False
    if "leaf_value" in a:
        return a["samples"] == b["samples"]
    return ((a["feature"], a["bin"]) == (b["feature"], b["bin"]) and _same_structure(a["left"], b["left"])
            and _same_structure(a["right"], b["right"]))


def _leaves(t):
    return [t["leaf_value"]] if "leaf_value" in t else _leaves(t["left"]) + _leaves(t["right"])


def test_fit_multi_regression_one_output_is_the_regression_fit():
    """M = 1 with unit weights reduces to the regression objective (pinned by
    its closed form above): same trees, leaves and scores."""
    _, codes, Y = _multi_data(21, M=1)
    eras = np.zeros(len(Y), int)
    base_m, recs_m, sc_m = O.fit_multi_regression(codes, Y, np.ones_like(Y), eras, 1, 16, 3, 6, 0.2)
    base_r, recs_r, sc_r = O.fit_regression(codes, Y[:, 0], np.ones(len(Y)), eras, 1, 16, 3, 6, 0.2)
    assert abs(base_m[0] - base_r) < 1e-12
    for a, b in zip(recs_m, recs_r):
        assert _same_structure(a.tree, b.tree)
        np.testing.assert_allclose([v[0] for v in _leaves(a.tree)], _leaves(b.tree), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(sc_m[:, 0], sc_r, rtol=1e-9, atol=1e-9)


def test_fit_multi_regression_unweighted_output_adds_nothing_to_the_splits():
    """An output with weight 0 everywhere contributes nothing to the summed
    split targets: the trees are output 0's regression trees, that output's
    leaves are 0 and its base is 0 (core.py:1061-1063)."""
    _, codes, Y = _multi_data(22, M=2)
    W = np.ones_like(Y)
    W[:, 1] = 0.0
    eras = np.zeros(len(Y), int)
    base, recs, sc = O.fit_multi_regression(codes, Y, W, eras, 1, 16, 3, 5, 0.3)
    _, recs_r, sc_r = O.fit_regression(codes, Y[:, 0], np.ones(len(Y)), eras, 1, 16, 3, 5, 0.3)
    assert base[1] == 0.0
    for a, b in zip(recs, recs_r):
        assert _same_structure(a.tree, b.tree)
        lv = np.array(_leaves(a.tree))
        np.testing.assert_allclose(lv[:, 0], _leaves(b.tree), rtol=1e-9, atol=1e-12)
        assert np.all(lv[:, 1] == 0.0)
    np.testing.assert_allclose(sc[:, 0], sc_r, rtol=1e-9, atol=1e-9)
    assert np.all(sc[:, 1] == 0.0)


def test_fit_multi_regression_missing_targets_are_zero_weights_and_base_brute_force():
    """A NaN target is the same as that entry's weight set to 0 (same trees,
    leaves, scores); the base is output by output the weighted mean of the
    finite targets, summed row by row."""
    rng, codes, Y = _multi_data(23, M=3)
    W = rng.uniform(0.5, 2.0, Y.shape)
    miss = rng.random(Y.shape) < 0.2
    Yn = Y.copy()
    Yn[miss] = np.nan
    Yn[0, 0] = np.inf
    miss[0, 0] = True
    eras = np.zeros(len(Y), int)
    b1, r1, s1 = O.fit_multi_regression(codes, Yn, W, eras, 1, 16, 3, 4, 0.3)
    b2, r2, s2 = O.fit_multi_regression(codes, np.where(miss, 0.0, Y), np.where(miss, 0.0, W), eras, 1, 16, 3, 4, 0.3)
    np.testing.assert_array_equal(b1, b2)
    for a, b in zip(r1, r2):
        assert _same_structure(a.tree, b.tree)
        np.testing.assert_array_equal(np.array(_leaves(a.tree)), np.array(_leaves(b.tree)))
    np.testing.assert_array_equal(s1, s2)
    for m in range(3):
        num = den = 0.0
        for i in range(len(Y)):
            if not miss[i, m]:
                num += W[i, m] * Y[i, m]
                den += W[i, m]
        assert abs(b1[m] - num / den) < 1e-12 * max(1.0, abs(num / den))


def test_fit_multi_regression_output_permutation_and_leaf_brute_force():
    """Permuting the outputs permutes every leaf vector and the scores; each
    leaf's value for output m equals sum g_m / (sum h_m + lambda) over the
    rows that reach it, summed directly."""
    rng, codes, Y = _multi_data(24, M=3)
    W = rng.uniform(0.5, 2.0, Y.shape)
    eras = np.zeros(len(Y), int)
    perm = [2, 0, 1]
    lam = 0.5
    b1, r1, s1 = O.fit_multi_regression(codes, Y, W, eras, 1, 16, 3, 3, 0.3, lam=lam)
    b2, r2, s2 = O.fit_multi_regression(codes, Y[:, perm], W[:, perm], eras, 1, 16, 3, 3, 0.3, lam=lam)
    np.testing.assert_allclose(b2, b1[perm], rtol=1e-12)
    np.testing.assert_allclose(s2, s1[:, perm], rtol=1e-9, atol=1e-9)
    for a, b in zip(r1, r2):
        assert _same_structure(a.tree, b.tree)
        np.testing.assert_allclose(np.array(_leaves(b.tree)), np.array(_leaves(a.tree))[:, perm], rtol=1e-9,
                                   atol=1e-12)
    # leaf brute force on the first tree, whose targets start from the base
    rows_of = _node_rows(r1[0].tree, codes)
    g = (Y - b1) * W
    for heap, node in r1[0].nodes.items():
        if node[0] != "leaf":
            continue
        rows = rows_of[heap]
        for m in range(3):
            ref = sum(g[i, m] for i in rows) / (sum(W[i, m] for i in rows) + lam)
            assert abs(node[1][m] - ref) < 1e-9 * max(1.0, abs(ref))
    np.testing.assert_allclose(O.predict_multi(codes, [r.tree for r in r1], 0.3, b1), s1, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- C restatement
def _c_tree_matches(ctab, rec_tree, rtol=1e-9):
    from oracle import cpu_gbm
    st, ft, bt, vt, ct = ctab
    d = cpu_gbm.to_dict(st, ft, bt, vt, ct)

    def cmp(a, b, path):
        assert ("leaf_value" in a) == ("leaf_value" in b), f"{path}: leaf/split differ"
        if "leaf_value" in a:
            assert a["samples"] == b["samples"], path
            assert abs(a["leaf_value"] - b["leaf_value"]) <= rtol * max(1.0, abs(b["leaf_value"])), path
            return
        assert (a["feature"], a["bin"]) == (int(b["feature"]), int(b["bin"])), path
        cmp(a["left"], b["left"], path + "L")
        cmp(a["right"], b["right"], path + "R")

    cmp(d, rec_tree, "root")


@pytest.mark.parametrize("E,mono", [(1, False), (1, True), (5, False), (5, True)])
def test_c_restatement_grows_the_oracle_tree(E, mono):
    """The multithreaded C tree growth (sibling histograms by subtraction,
    thread-private sums) gives the NumPy oracle's tree: same splits, row
    counts and leaf values, and the same row -> leaf map."""
    from oracle import cpu_gbm
    rng = np.random.default_rng(40 + E)
    N, F, B, D = 6000, 10, 16, 5
    codes = rng.integers(0, B, (N, F)).astype(np.int8)
    g = np.sin(codes[:, 0] * 0.7) + 0.5 * (codes[:, 3] > 7) + 0.3 * rng.standard_normal(N)
    h = rng.uniform(0.5, 1.5, N)
    eras = rng.integers(0, E, N).astype(np.int32)
    req = np.zeros(F, np.int8)
    if mono:
        req[0], req[3] = 1, -1
    rec = O.grow_tree(codes, g, h, eras, E, B, D, 20, 0.0, 1e-6, mono_required=req if mono else None)
    ctab, leaf = cpu_gbm.grow_tree(codes, g, h, eras, E, B, D, 20, 0.0, 1e-6, mono_required=req if mono else None,
                                   threads=4)
    _c_tree_matches(ctab, rec.tree)
    np.testing.assert_array_equal(leaf, rec.leaf_of_row)


def test_c_restatement_boosting_loops_match_the_oracle():
    """Regression boosting equals fit_regression (scores within 1e-9); the
    first softmax round's K trees equal the oracle trees grown on
    softmax_grad_hess of the log-prior scores."""
    from oracle import cpu_gbm
    rng = np.random.default_rng(44)
    N, F, B = 4000, 8, 32
    codes = rng.integers(0, B, (N, F)).astype(np.int8)
    y = codes[:, 0] * 0.1 - (codes[:, 2] > 20) + 0.2 * rng.standard_normal(N)
    w = rng.uniform(0.5, 2.0, N)
    eras = np.zeros(N, np.int32)
    base_o, recs, sc_o = O.fit_regression(codes, y, w, eras, 1, B, 4, 5, 0.3)
    base_c, tabs, sc_c = cpu_gbm.fit_regression(codes, y, w, eras, 1, B, 4, 5, 0.3, threads=3)
    assert abs(base_c - base_o) < 1e-12
    for t, rec in enumerate(recs):
        _c_tree_matches(tuple(a[t] for a in tabs), rec.tree)
    np.testing.assert_allclose(sc_c, sc_o, rtol=1e-9, atol=1e-9)
    K = 3
    lab = (codes[:, 1] % K).astype(np.int32)
    tabs_k, _ = cpu_gbm.fit_softmax(codes, lab, np.ones(N), K, B, 3, 1, 0.1, threads=2)
    counts = np.bincount(lab, minlength=K)
    gk, hk = O.softmax_grad_hess(np.tile(np.log(counts / N + 1e-10), (N, 1)), lab, np.ones(N))
    for k in range(K):
        rec = O.grow_tree(codes, gk[:, k], hk[:, k], eras, 1, B, 3, 20, 0.0, 1e-6)
        _c_tree_matches(tuple(a[0, k] for a in tabs_k), rec.tree)
