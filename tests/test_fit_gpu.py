"""End-to-end GPU parity of WarpGBM.fit/predict against the oracle.

* plain per-node engine and batched level-wise engine give identical trees
  (same split at every node, same leaf values bit for bit);
* each tree matches the FP64 oracle's tree grown from the same float32
  gradients (teacher forcing): same (feature, bin) at every node, same
  row -> leaf assignment, leaf values within 1e-5 relative;
* predictions within 1e-5 of the oracle's float64 forest evaluation.
"""
import numpy as np
import pytest

from oracle import gbm_oracle as O
import _parity as PAR

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _data_cfg1(n=20000, f=50, seed=0):
    """Plain regression, small (20k x 50, y = X[:, :10] w + noise)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f)).astype(np.float32)
    w = rng.standard_normal(10)
    y = (X[:, :10] @ w + rng.normal(0, 0.1, n)).astype(np.float32)
    return X, y


def _trees_equal(a, b, path="root"):
    if ("leaf_value" in a) != ("leaf_value" in b):
        return f"{path}: leaf/split mismatch"
    if "leaf_value" in a:
        if a["leaf_value"] != b["leaf_value"] or a["samples"] != b["samples"]:
            return f"{path}: leaf {a} vs {b}"
        return None
    fa = int(a["feature"].item()) if torch.is_tensor(a["feature"]) else int(a["feature"])
    fb = int(b["feature"].item()) if torch.is_tensor(b["feature"]) else int(b["feature"])
    if fa != fb or a["bin"] != b["bin"]:
        return f"{path}: split ({fa},{a['bin']}) vs ({fb},{b['bin']})"
    return _trees_equal(a["left"], b["left"], path + "L") or _trees_equal(a["right"], b["right"], path + "R")


def _oracle_check_tree(tree, codes, g, h, eras, E, B, D, lam):
    """Grow the oracle tree from the same g/h and compare node by node."""
    rec = O.grow_tree(codes, g, h, eras, E, B, D, 20, 0.0, lam)
    msg = _compare_dict_to_oracle(tree, rec.tree, "root")
    assert msg is None, msg


def _compare_dict_to_oracle(t, o, path):
    if "leaf_value" in o:
        assert "leaf_value" in t, f"{path}: engine split where oracle has a leaf"
        assert t["samples"] == o["samples"], f"{path}: sample count {t['samples']} vs {o['samples']}"
        ref = o["leaf_value"]
        assert abs(t["leaf_value"] - ref) <= 1e-5 * max(abs(ref), 1e-3), f"{path}: leaf {t['leaf_value']} vs {ref}"
        return None
    assert "feature" in t, f"{path}: engine leaf where oracle splits"
    f = int(t["feature"].item())
    assert (f, t["bin"]) == (o["feature"], o["bin"]), f"{path}: split ({f},{t['bin']}) vs ({o['feature']},{o['bin']})"
    _compare_dict_to_oracle(t["left"], o["left"], path + "L")
    _compare_dict_to_oracle(t["right"], o["right"], path + "R")
    return None


def test_cfg1_plain_equals_batched_and_oracle():
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1()
    kw = dict(objective="regression", num_bins=32, max_depth=4, n_estimators=20, learning_rate=0.1, random_state=0)
    mb = WarpGBM(engine="batched", **kw).fit(X, y)
    mp = WarpGBM(engine="plain", **kw).fit(X, y)
    for i in range(20):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    pb, pp = mb.predict(X), mp.predict(X)
    assert np.array_equal(pb, pp)
    np.testing.assert_array_equal(mb.feature_importance_, mp.feature_importance_)

    # oracle with teacher forcing: replay the engine's float32 scores
    codes, _ = O.bin_matrix(X, 32)
    eras = np.zeros(len(y), np.int32)
    scores = np.full(len(y), np.float32(mb.base_prediction), np.float32)
    lr = np.float32(0.1)
    flats = []
    for i in range(20):
        g = ((y - scores) * np.float32(1.0)).astype(np.float32)
        h = np.ones_like(g)
        _oracle_check_tree(mb.forest[i], codes, g, h, eras, 1, 32, 4, 1e-6)
        vals = O.predict_tree(mb.forest[i] if False else _to_plain(mb.forest[i]), codes)
        scores = (scores + (lr * vals.astype(np.float32)).astype(np.float32)).astype(np.float32)
        flats.append(O.flatten_tree(_to_plain(mb.forest[i]), 2 ** 5))
    ref = O.predict_forest_flat(codes, flats, 0.1, base=mb.base_prediction)
    np.testing.assert_allclose(pb, ref, rtol=1e-5, atol=1e-5)
    # training scores and predictions of the training matrix coincide bit for bit
    assert np.array_equal(pb, scores)


def _to_plain(t):
    if "leaf_value" in t:
        return dict(t)
    return {"feature": int(t["feature"].item()), "bin": t["bin"], "left": _to_plain(t["left"]),
            "right": _to_plain(t["right"])}


def test_eras_and_monotone_batched_equals_plain():
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(5)
    n, f = 12000, 12
    X = rng.standard_normal((n, f)).astype(np.float32)
    era = np.repeat(np.arange(6), n // 6)
    rng.shuffle(era)
    sign = np.where(era % 3 == 0, -1.0, 1.0)
    y = (X[:, 0] * sign + 0.5 * X[:, 1] + 0.1 * rng.standard_normal(n)).astype(np.float32)
    kw = dict(objective="regression", num_bins=16, max_depth=3, n_estimators=8, learning_rate=0.2,
              monotonic_constraints={1: 1})
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(8):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)
    assert np.array_equal(mb.predict(X), mp.predict(X))


def test_conflicting_eras_root_is_leaf():
    """warpgbm tests/test_monotonic_constraints.py:124-149."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(90)
    n = 800
    x0 = rng.uniform(-2.0, 2.0, size=n).astype(np.float32)
    era = np.zeros(n, dtype=np.int32)
    era[n // 2:] = 1
    y = np.empty(n, dtype=np.float32)
    y[: n // 2] = 2.0 * x0[: n // 2]
    y[n // 2:] = -2.0 * x0[n // 2:]
    y += 0.01 * rng.normal(size=n).astype(np.float32)
    for engine in ("batched", "plain"):
        m = WarpGBM(objective="regression", n_estimators=10, max_depth=3, learning_rate=0.1, num_bins=32,
                    monotonic_constraints=[1], random_state=90, engine=engine)
        m.fit(x0.reshape(-1, 1), y, era_id=era)
        assert "leaf_value" in m.forest[0]


def test_multiclass_batched_equals_plain():
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(7)
    n, f, K = 6000, 20, 4
    X = rng.standard_normal((n, f)).astype(np.float32)
    W = rng.standard_normal((8, K))
    lab = np.argmax(X[:, :8] @ W + rng.gumbel(size=(n, K)), axis=1)
    kw = dict(objective="multiclass", num_bins=32, max_depth=3, n_estimators=6, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, lab)
    mp = WarpGBM(engine="plain", **kw).fit(X, lab)
    for i in range(6):
        for k in range(K):
            msg = _trees_equal(mb.forest[i][k], mp.forest[i][k], f"round{i} class{k}:")
            assert msg is None, msg
    pb = mb.predict_proba(X)
    assert np.array_equal(pb, mp.predict_proba(X))
    np.testing.assert_allclose(pb.sum(axis=1), 1.0, atol=1e-5)


def test_graph_replay_equals_eager():
    """Rounds replayed from one captured graph (device-side round counter,
    per-round feature subsets) give the trees of the eager launch sequence;
    histogram timing forces the eager path."""
    import ctypes
    from warpgbm_amd import WarpGBM, _lib
    X, y = _data_cfg1(n=8000, f=30, seed=3)
    kw = dict(objective="regression", num_bins=32, max_depth=4, n_estimators=7, learning_rate=0.1,
              colsample_bytree=0.6)
    torch.manual_seed(11)
    mg = WarpGBM(engine="batched", **kw).fit(X, y)
    assert _lib.query("wgbm_fit_graph_rounds") == 6
    _lib.call("wgbm_profile_begin")
    try:
        torch.manual_seed(11)
        me = WarpGBM(engine="batched", **kw).fit(X, y)
        assert _lib.query("wgbm_fit_graph_rounds") == 0
    finally:
        _lib.call("wgbm_profile_end", (ctypes.c_double * 4)())
    for i in range(7):
        msg = _trees_equal(mg.forest[i], me.forest[i], f"tree{i}:")
        assert msg is None, msg
    assert np.array_equal(mg.predict(X), me.predict(X))


def test_many_eras_prebinned_batched_equals_plain_and_oracle():
    """Pre-binned int8 codes in [0, 5) with 40 eras of uneven size (the
    many-era path: per-era gains, ordered era sums): both engines agree and
    the first tree matches the oracle grown from the same float32 targets."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(21)
    n, f, E = 16000, 24, 40
    X = rng.integers(0, 5, (n, f)).astype(np.int8)
    sizes = rng.lognormal(5.0, 0.5, E)
    sizes = np.floor(sizes / sizes.sum() * n).astype(int)
    sizes[np.argmax(sizes)] += n - sizes.sum()
    era = np.repeat(np.arange(E), sizes).astype(np.int32)
    flip = np.where(rng.random(E) < 0.4, -1.0, 1.0)
    sig = (X[:, 0] - 2.0) * flip[era] + 0.7 * (X[:, 1] - 2.0) + 0.3 * rng.standard_normal(n)
    y = (np.searchsorted(np.quantile(sig, [0.2, 0.4, 0.6, 0.8]), sig) / 4.0).astype(np.float32)
    kw = dict(objective="regression", num_bins=5, max_depth=5, n_estimators=6, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(6):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    _oracle_check_tree(mb.forest[0], X, g, np.ones_like(g), era, E, 5, 5, 1e-6)


def test_deep_levels_of_many_eras_match_oracle():
    """400 eras at depth 8: levels 5 and 6 have 12800 and 25600 (node, era)
    segments and ~10k partition chunks, so the level's plans and partition
    scan run their multi-pass forms (staged scans of 4096 values, gather
    loops of several rounds), and the 128-node level selects its splits in
    the fused per-era gain kernel.  The first tree matches the oracle and
    both engines agree on every tree."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(33)
    n, f, E = 240000, 8, 400
    X = rng.integers(0, 5, (n, f)).astype(np.int8)
    sizes = rng.lognormal(5.0, 0.4, E)
    sizes = np.floor(sizes / sizes.sum() * n).astype(int)
    sizes[np.argmax(sizes)] += n - sizes.sum()
    era = np.repeat(np.arange(E), sizes).astype(np.int32)
    w = np.array([1.0, 0.9, 0.8, 0.7, 0.6, 0.5, 0.45, 0.4], np.float32)
    sig = X[:, :8].astype(np.float32) @ w + 0.2 * rng.standard_normal(n).astype(np.float32)
    y = (sig / 8.0).astype(np.float32)
    kw = dict(objective="regression", num_bins=5, max_depth=8, n_estimators=3, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(3):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    _oracle_check_tree(mb.forest[0], X, g, np.ones_like(g), era, E, 5, 8, 1e-6)
    # monotone constraints on the same levels: every era's direction must
    # match (the fused kernel's mask), against the per-node selection op
    kw_m = dict(kw, n_estimators=2, monotonic_constraints={0: 1, 2: -1, 5: 1})
    mb = WarpGBM(engine="batched", **kw_m).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw_m).fit(X, y, era_id=era)
    for i in range(2):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"monotone tree{i}:")
        assert msg is None, msg


def test_multiclass_round0_matches_oracle():
    """Softmax targets of the first round (device float32) match the FP64
    oracle within 1e-6, and every class tree of round 0 matches the oracle
    tree grown from the same float32 targets."""
    from warpgbm_amd import WarpGBM, _lib
    rng = np.random.default_rng(8)
    n, f, K = 5000, 12, 4
    X = rng.standard_normal((n, f)).astype(np.float32)
    lab = np.argmax(X[:, :6] @ rng.standard_normal((6, K)) + rng.gumbel(size=(n, K)), axis=1)
    m = WarpGBM(objective="multiclass", num_bins=32, max_depth=3, n_estimators=1, learning_rate=0.1).fit(X, lab)
    codes, _ = O.bin_matrix(X, 32)
    counts = np.bincount(lab, minlength=K)
    sc = np.empty((n, K), np.float32)
    for k in range(K):
        sc[:, k] = np.float32(np.log(np.float32(np.float32(counts[k]) / np.float32(n)) + np.float32(1e-10)))
    dev = torch.device("cuda")
    s_d = torch.from_numpy(sc).to(dev)
    l_d = torch.from_numpy(lab.astype(np.int32)).to(dev)
    w_d = torch.ones(n, dtype=torch.float32, device=dev)
    g_d = torch.empty(K * n, dtype=torch.float32, device=dev)
    h_d = torch.empty(K * n, dtype=torch.float32, device=dev)
    _lib.call("wgbm_grad_softmax", s_d.data_ptr(), l_d.data_ptr(), w_d.data_ptr(), n, K, g_d.data_ptr(),
              h_d.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    g = g_d.cpu().numpy().reshape(K, n)
    h = h_d.cpu().numpy().reshape(K, n)
    g64, h64 = O.softmax_grad_hess(sc.astype(np.float64), lab, np.ones(n))
    np.testing.assert_allclose(g, g64.T, atol=1e-6)
    np.testing.assert_allclose(h, h64.T, atol=1e-6)
    eras = np.zeros(n, np.int32)
    for k in range(K):
        _oracle_check_tree(m.forest[0][k], codes, g[k], h[k], eras, 1, 32, 3, 1e-6)


def test_benchmark_size_trees_match_plain_and_oracle_argmax():
    """The benchmark's configuration at its full size (1M x 1000 ~N(0,1),
    y = tanh(X[:, :20] w) + noise, 32 quantile bins, depth 5), 3 trees from
    device-resident codes: the batched engine's trees equal the plain
    engine's, and every split of the first tree is an argmax of the FP64
    oracle's gains over all 1000 x 31 thresholds, with histograms counted
    from the same codes and gradients over the node's rows (the engine's
    gain at its split within 1e-9 relative of the oracle's best, so a float
    near-tie cannot fail it); its leaves hold exactly the rows its splits
    send there, with FP64 leaf values within 1e-5."""
    from warpgbm_amd import WarpGBM
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    N, F, B = 1_000_000, 1000, 32
    X = torch.randn((N, F), generator=gen, device=dev, dtype=torch.float32)
    w = torch.randn((20,), generator=gen, device=dev, dtype=torch.float32)
    y = (torch.tanh(X[:, :20] @ w) + 0.1 * torch.randn((N,), generator=gen, device=dev)).contiguous()
    m0 = WarpGBM(num_bins=B)
    m0.device = dev
    edges = m0._quantile_edges(X)
    codes = m0._bin_matrix(X, edges)
    del X
    torch.cuda.empty_cache()
    kw = dict(objective="regression", num_bins=B, max_depth=5, n_estimators=3, learning_rate=0.01, device=dev)
    mb = WarpGBM(engine="batched", **kw)
    mb.fit_binned(codes, y, bin_edges=edges)
    mp = WarpGBM(engine="plain", **kw)
    mp.fit_binned(codes, y, bin_edges=edges)
    for i in range(3):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg

    cols = codes.t().contiguous().cpu().numpy()  # [F, N]
    del codes
    yh = y.cpu().numpy()
    g = (yh - np.float32(mb.base_prediction)).astype(np.float64)

    worst_leaf = [0.0]

    def check(node, rows, path, depth):
        """Every node of tree 0, following the engine's own partition: each
        split is an argmax of the FP64 gains over the node's rows, each leaf
        holds exactly the rows the splits send there (row -> leaf
        assignment) with the FP64 leaf value."""
        gr = g[rows]
        if "leaf_value" in node:
            assert node["samples"] == rows.size, f"{path}: {node['samples']} rows vs {rows.size}"
            ref = gr.sum() / (rows.size + 1e-6)
            err = abs(node["leaf_value"] - ref) / max(abs(ref), 1e-3)
            worst_leaf[0] = max(worst_leaf[0], err)
            assert err <= 1e-5, f"{path}: leaf {node['leaf_value']} vs {ref}"
            return
        assert depth < 5
        G = np.zeros((1, F, B))
        H = np.zeros((1, F, B))
        for f in range(F):
            c = cols[f][rows]
            G[0, f] = np.bincount(c, weights=gr, minlength=B)
            H[0, f] = np.bincount(c, minlength=B)
        gain, direction = O.split_gains(G, H, 20.0, 1e-6)
        of, ob = O.select_split(gain, direction, 0.0)
        assert of >= 0, f"{path}: oracle finds no split"
        f, b = int(node["feature"].item()), node["bin"]
        best = gain[0, of, ob]
        assert gain[0, f, b] >= best - 1e-9 * abs(best), f"{path}: ({f},{b}) gain {gain[0, f, b]} < oracle ({of},{ob}) {best}"
        go_left = cols[f][rows] <= b
        check(node["left"], rows[go_left], path + "L", depth + 1)
        check(node["right"], rows[~go_left], path + "R", depth + 1)

    check(mb.forest[0], np.arange(N), "root", 0)
    PAR.record("1M x 1000 tree 0 leaf values (relative)", worst_leaf[0], 1e-5)


def test_skewed_eras_multi_chunk_segments_match_oracle():
    """1M rows in 200 eras, one era holding 95 % of them, depth 6: levels of
    3200 and 6400 (node, era) segments whose giant-era segments span 29-58
    partition chunks and several histogram chunks, so the blocked plan
    passes write multi-chunk segments both per thread and by the whole
    workgroup.  Both engines agree and the first tree matches the oracle."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(8)
    n, f, E = 1_000_000, 8, 200
    X = rng.integers(0, 5, (n, f)).astype(np.int8)
    big = int(0.95 * n)
    sizes = np.full(E - 1, (n - big) // (E - 1))
    sizes[-1] += n - big - sizes.sum()
    era = np.concatenate([np.zeros(big, np.int32), np.repeat(np.arange(1, E), sizes).astype(np.int32)])
    w = np.array([1.0, 0.9, 0.8, 0.7, 0.6, 0.5, 0.45, 0.4], np.float32)
    y = ((X.astype(np.float32) @ w + 0.2 * rng.standard_normal(n).astype(np.float32)) / 8.0).astype(np.float32)
    kw = dict(objective="regression", num_bins=5, max_depth=6, n_estimators=2, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(2):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    _oracle_check_tree(mb.forest[0], X, g, np.ones_like(g), era, E, 5, 6, 1e-6)


def test_twenty_million_rows_match_plain_and_oracle():
    """20M rows x 4 pre-binned features, 8 bins, depth 3: root bins of ~2.5M
    rows, ~20k partition chunks per level (the multi-pass scans), histogram
    segments of thousands of chunks.  Both engines agree and the first tree
    matches the oracle grown from the same gradients."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(12)
    n, f, B = 20_000_000, 4, 8
    X = rng.integers(0, B, (n, f)).astype(np.int8)
    y = (X.astype(np.float32) @ np.array([1.0, -0.7, 0.5, 0.2], np.float32)
         + rng.standard_normal(n).astype(np.float32)).astype(np.float32)
    kw = dict(objective="regression", num_bins=B, max_depth=3, n_estimators=2, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y)
    mp = WarpGBM(engine="plain", **kw).fit(X, y)
    for i in range(2):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    _oracle_check_tree(mb.forest[0], X, g, np.ones_like(g), np.zeros(n, np.int32), 1, B, 3, 1e-6)


def test_thousands_of_eras_match_plain_and_oracle():
    """3000 eras of ~20 rows (60k rows), depth 3: the root alone has 3000
    (node, era) segments, so even the root's plans take the blocked passes,
    and the era count / scan / scatter kernels run several passes.  Both
    engines agree and the first tree matches the oracle."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(31)
    n, f, E = 60000, 6, 3000
    X = rng.integers(0, 5, (n, f)).astype(np.int8)
    era = rng.integers(0, E, n).astype(np.int32)
    y = ((X[:, 0] - 2.0) + 0.5 * (X[:, 1] - 2.0) + 0.3 * rng.standard_normal(n)).astype(np.float32)
    kw = dict(objective="regression", num_bins=5, max_depth=3, n_estimators=2, learning_rate=0.1,
              min_child_weight=1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(2):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    rec = O.grow_tree(X, g, np.ones_like(g), era, E, 5, 3, 1, 0.0, 1e-6)
    msg = _compare_dict_to_oracle(mb.forest[0], rec.tree, "root")
    assert msg is None, msg


@pytest.mark.parametrize("B", [2, 127])
def test_bin_count_extremes_match_plain_and_oracle(B):
    """The smallest and largest bin counts (2 and 127): quantile binning,
    one-feature-per-lane histograms of 128 spare-padded bins, the one-feature
    wave scan; both engines agree and every tree matches the oracle under
    teacher forcing."""
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1(n=12000, f=20, seed=B)
    kw = dict(objective="regression", num_bins=B, max_depth=4, n_estimators=4, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y)
    mp = WarpGBM(engine="plain", **kw).fit(X, y)
    for i in range(4):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    codes, _ = O.bin_matrix(X, B)
    eras = np.zeros(len(y), np.int32)
    scores = np.full(len(y), np.float32(mb.base_prediction), np.float32)
    for i in range(4):
        g = (y - scores).astype(np.float32)
        _oracle_check_tree(mb.forest[i], codes, g, np.ones_like(g), eras, 1, B, 4, 1e-6)
        vals = O.predict_tree(_to_plain(mb.forest[i]), codes)
        scores = (scores + (np.float32(0.1) * vals.astype(np.float32)).astype(np.float32)).astype(np.float32)


def test_multiclass_with_eras_and_shallow_trees_match_plain():
    """Softmax classes and eras together (slots = class x node x era), depth 1
    stumps and a node too small to split: both engines grow the same trees
    and give the same probabilities."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(17)
    n, f, K, E = 8000, 10, 3, 5
    X = rng.standard_normal((n, f)).astype(np.float32)
    era = rng.integers(0, E, n).astype(np.int32)
    lab = np.argmax(X[:, :5] @ rng.standard_normal((5, K)) + rng.gumbel(size=(n, K)), axis=1)
    for depth in (1, 3):
        kw = dict(objective="multiclass", num_bins=16, max_depth=depth, n_estimators=3, learning_rate=0.2)
        mb = WarpGBM(engine="batched", **kw).fit(X, lab, era_id=era)
        mp = WarpGBM(engine="plain", **kw).fit(X, lab, era_id=era)
        for i in range(3):
            for k in range(K):
                msg = _trees_equal(mb.forest[i][k], mp.forest[i][k], f"depth{depth} round{i} class{k}:")
                assert msg is None, msg
        np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)
        assert np.array_equal(mb.predict_proba(X), mp.predict_proba(X))
    # fewer rows than two minimum children: every tree is a single leaf
    m = WarpGBM(objective="regression", num_bins=8, max_depth=3, n_estimators=2, engine="batched")
    m.fit(X[:30], X[:30, 0].copy())
    assert all("leaf_value" in t and t["samples"] == 30 for t in m.forest)


def test_depth_beyond_the_level_tables_runs_per_node():
    """max_depth 22 exceeds the batched engine's level tables (depth <= 20):
    the fit runs on the per-node engine and grows the oracle's tree (growth
    stops where children would fall below min_child_weight)."""
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1(n=3000, f=12, seed=22)
    m = WarpGBM(objective="regression", num_bins=16, max_depth=22, n_estimators=1, learning_rate=0.1)
    m.fit(X, y)
    codes, _ = O.bin_matrix(X, 16)
    g = (y - np.float32(m.base_prediction)).astype(np.float32)
    _oracle_check_tree(m.forest[0], codes, g, np.ones_like(g), np.zeros(len(y), np.int32), 1, 16, 22, 1e-6)


def test_column_sampling_with_eras_batched_equals_plain():
    """colsample_bytree < 1 with eras: each round's feature subset (the same
    seeded draws in both engines) indexes the per-era histograms, gains and
    importances; the engines grow identical trees."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(41)
    n, f, E = 9000, 16, 3
    X = rng.standard_normal((n, f)).astype(np.float32)
    era = rng.integers(0, E, n).astype(np.int32)
    y = (X[:, 0] + 0.7 * X[:, 3] - 0.5 * X[:, 7] + 0.2 * rng.standard_normal(n)).astype(np.float32)
    kw = dict(objective="regression", num_bins=16, max_depth=3, n_estimators=6, learning_rate=0.2,
              colsample_bytree=0.5, random_state=5)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(6):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)


def test_weights_many_classes_and_wide_sampling_batched_equals_plain():
    """Sample weights with softmax (two hessian planes) and with eras;
    20 classes (20 trees per round in one level launch); 1000 features
    sampled at 30 % per round: both engines grow identical trees."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(53)
    n, f = 6000, 12
    X = rng.standard_normal((n, f)).astype(np.float32)
    w = rng.uniform(0.2, 3.0, n).astype(np.float32)
    era = rng.integers(0, 4, n).astype(np.int32)
    cases = []
    lab3 = np.argmax(X[:, :4] @ rng.standard_normal((4, 3)) + rng.gumbel(size=(n, 3)), axis=1)
    cases.append((dict(objective="multiclass", num_bins=16, max_depth=3, n_estimators=3), X, lab3,
                  dict(sample_weight=w)))
    y = (X[:, 0] - 0.5 * X[:, 2] + 0.2 * rng.standard_normal(n)).astype(np.float32)
    cases.append((dict(objective="regression", num_bins=16, max_depth=3, n_estimators=3), X, y,
                  dict(sample_weight=w, era_id=era)))
    lab20 = np.argmax(X[:, :6] @ rng.standard_normal((6, 20)) + rng.gumbel(size=(n, 20)), axis=1)
    cases.append((dict(objective="multiclass", num_bins=8, max_depth=2, n_estimators=2), X, lab20, {}))
    Xw = rng.standard_normal((3000, 1000)).astype(np.float32)
    yw = (Xw[:, 5] + Xw[:, 500] - Xw[:, 999] + 0.1 * rng.standard_normal(3000)).astype(np.float32)
    cases.append((dict(objective="regression", num_bins=8, max_depth=3, n_estimators=4, colsample_bytree=0.3,
                       random_state=3), Xw, yw, {}))
    for ci, (kw, Xc, yc, fk) in enumerate(cases):
        mb = WarpGBM(engine="batched", learning_rate=0.2, **kw).fit(Xc, yc, **fk)
        mp = WarpGBM(engine="plain", learning_rate=0.2, **kw).fit(Xc, yc, **fk)
        for i in range(kw["n_estimators"]):
            tb, tp = mb.forest[i], mp.forest[i]
            pairs = zip(tb, tp) if kw["objective"] == "multiclass" else [(tb, tp)]
            for k, (a, b) in enumerate(pairs):
                msg = _trees_equal(a, b, f"case{ci} tree{i} class{k}:")
                assert msg is None, msg


def test_missing_values_bin_and_predict_like_the_oracle():
    """Float inputs with NaN: a column holding any NaN, even one, gets NaN
    quantile edges (as torch.quantile gives) and every value of it code 0.
    Device codes equal the oracle's bit for bit, both engines grow the same
    trees and predictions match the oracle's forest on the same codes."""
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1(n=6000, f=12, seed=61)
    rng = np.random.default_rng(61)
    X[rng.random(len(X)) < 0.02, 3] = np.nan
    X[17, 5] = np.nan
    kw = dict(objective="regression", num_bins=16, max_depth=3, n_estimators=3, learning_rate=0.1)
    mb = WarpGBM(engine="batched", **kw).fit(X, y)
    mp = WarpGBM(engine="plain", **kw).fit(X, y)
    for i in range(3):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    codes, edges = O.bin_matrix(X, 16)
    for col in (3, 5):
        assert np.isnan(edges[col]).all() and (codes[:, col] == 0).all()
    be = mb.bin_edges.cpu().numpy() if torch.is_tensor(mb.bin_edges) else np.asarray(mb.bin_edges)
    np.testing.assert_array_equal(be, edges)  # NaN where the oracle has NaN
    Xd = torch.from_numpy(X).to("cuda")
    np.testing.assert_array_equal(mb._bin_matrix(Xd, torch.from_numpy(edges).to("cuda")).cpu().numpy(), codes)
    flats = [O.flatten_tree(_to_plain(t), 2 ** 4) for t in mb.forest]
    ref = O.predict_forest_flat(codes, flats, 0.1, base=mb.base_prediction)
    np.testing.assert_allclose(mb.predict(X), ref, rtol=1e-5, atol=1e-5)


def test_random_configurations_batched_equals_plain_and_oracle():
    """24 seeded random configurations (objective, rows, features, bins 2..127,
    depth 1..6, eras, weights, column sampling, monotone constraints,
    min_child_weight, float or pre-binned inputs): both engines grow the
    same trees, and unweighted regression without sampling or constraints
    grows the oracle's first tree."""
    import contextlib
    import io
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(2024)
    for ci in range(24):
        obj = "multiclass" if rng.random() < 0.25 else "regression"
        N, F = int(rng.integers(200, 8000)), int(rng.integers(1, 30))
        B = int(rng.choice([2, 3, 5, 8, 16, 31, 32, 64, 127]))
        D, E = int(rng.integers(1, 7)), int(rng.choice([1, 1, 2, 3, 7, 40]))
        K = int(rng.integers(2, 6)) if obj == "multiclass" else 1
        pre = rng.random() < 0.3
        X = rng.integers(0, B, (N, F)).astype(np.int8) if pre else rng.standard_normal((N, F)).astype(np.float32)
        era = rng.integers(0, E, N).astype(np.int32)
        w = rng.uniform(0.1, 3.0, N).astype(np.float32) if rng.random() < 0.3 else None
        cs = float(rng.choice([1.0, 1.0, 0.5]))
        mono = {0: int(rng.choice([-1, 1]))} if obj == "regression" and rng.random() < 0.2 else None
        mcw = int(rng.choice([1, 5, 20]))
        Xf = X.astype(np.float32)
        if obj == "regression":
            y = (Xf[:, 0] - 0.5 * Xf[:, min(1, F - 1)] + 0.3 * rng.standard_normal(N)).astype(np.float32)
        else:
            y = np.arange(N) % K  # every class present
            rng.shuffle(y)
        kw = dict(objective=obj, num_bins=B, max_depth=D, n_estimators=2, learning_rate=0.2, colsample_bytree=cs,
                  random_state=ci, min_child_weight=mcw)
        if mono:
            kw["monotonic_constraints"] = mono
        fk = dict(era_id=era) if w is None else dict(era_id=era, sample_weight=w)
        with contextlib.redirect_stdout(io.StringIO()):
            mb = WarpGBM(engine="batched", **kw).fit(X, y, **fk)
            mp = WarpGBM(engine="plain", **kw).fit(X, y, **fk)
        for i in range(2):
            pairs = zip(mb.forest[i], mp.forest[i]) if obj == "multiclass" else [(mb.forest[i], mp.forest[i])]
            for k, (a, b) in enumerate(pairs):
                msg = _trees_equal(a, b, f"case {ci} ({obj} N={N} F={F} B={B} D={D} E={E}) tree{i} class{k}:")
                assert msg is None, msg
        if obj == "regression" and w is None and cs == 1.0 and mono is None:
            codes, _ = O.bin_matrix(X, B)
            g = (y - np.float32(mb.base_prediction)).astype(np.float32)
            eu = np.unique(era, return_inverse=True)[1].astype(np.int32)
            rec = O.grow_tree(codes, g, np.ones_like(g), eu, int(eu.max()) + 1, B, D, mcw, 0.0, 1e-6)
            msg = _compare_dict_to_oracle(mb.forest[0], rec.tree, f"case {ci} root")
            assert msg is None, msg


@pytest.mark.parametrize("weighted", [False, True])
def test_high_dynamic_range_residuals_batched_equals_plain_and_oracle(weighted):
    """95 % of the rows carry residuals near 1e-6, 5 % near +-1e3 (balanced,
    so the base is ~0): nodes holding both take the count-packed words with
    the exact fine path (unit weights) or 49-bit words (weights), in the
    per-node op and in the batched level builds alike.  Both engines grow
    the same trees, and tree 0 equals the oracle's grown from the same
    float32 targets."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(61)
    n, f = 30000, 10
    X = rng.standard_normal((n, f)).astype(np.float32)
    big = rng.random(n) < 0.05
    y = (1e-6 * (X[:, 1] + 0.3 * rng.standard_normal(n))).astype(np.float32)
    sgn = np.where(np.arange(big.sum()) % 2 == 0, 1.0, -1.0)
    y[big] = (1e3 * sgn * (1.0 + 0.1 * (X[big, 0] > 0))).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32) if weighted else None
    kw = dict(objective="regression", num_bins=16, max_depth=5, n_estimators=4, learning_rate=0.5,
              min_child_weight=5)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, sample_weight=w)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, sample_weight=w)
    for i in range(4):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    codes, _ = O.bin_matrix(X, 16)
    ww = np.ones(n, np.float32) if w is None else mb.sample_weight_gpu.cpu().numpy()
    g = ((y - np.float32(mb.base_prediction)) * ww).astype(np.float32)
    rec = O.grow_tree(codes, g, ww.astype(np.float64), np.zeros(n, np.int32), 1, 16, 5, 5, 0.0, 1e-6)
    msg = _compare_dict_to_oracle(mb.forest[0], rec.tree, "root")
    assert msg is None, msg


@pytest.mark.parametrize("depth,K", [(11, 1), (8, 5)])
def test_fused_levels_then_stable_levels_batched_equals_plain(depth, K):
    """One-era levels of at most 512 nodes take the fused single-pass split;
    deeper levels (1024 nodes and more) the stable partition.  A tree that
    crosses from one to the other (depth 11; 5 classes at depth 8: levels of
    5 * 2^d nodes) grows the per-node engine's trees, leaf values and scores
    bit for bit, and its first tree is the oracle's."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(40 + depth)
    n, f = 40000, 16
    X = rng.standard_normal((n, f)).astype(np.float32)
    if K == 1:
        y = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.1 * rng.standard_normal(n)).astype(np.float32)
        kw = dict(objective="regression")
    else:
        y = np.argmax(X[:, :K] + 0.3 * rng.standard_normal((n, K)), axis=1)
        kw = dict(objective="multiclass")
    kw.update(num_bins=32, max_depth=depth, n_estimators=3, learning_rate=0.2)
    mb = WarpGBM(engine="batched", **kw).fit(X, y)
    mp = WarpGBM(engine="plain", **kw).fit(X, y)
    for i in range(3):
        ta = mb.forest[i] if K == 1 else mb.forest[i]
        tb = mp.forest[i] if K == 1 else mp.forest[i]
        if K == 1:
            msg = _trees_equal(ta, tb, f"tree{i}:")
            assert msg is None, msg
        else:
            for c in range(K):
                msg = _trees_equal(ta[c], tb[c], f"round{i} class{c}:")
                assert msg is None, msg
    if K == 1:
        assert np.array_equal(mb.predict(X), mp.predict(X))
        codes, _ = O.bin_matrix(X, 32)
        g = (y - np.float32(mb.base_prediction)).astype(np.float32)
        rec = O.grow_tree(codes, g, np.ones_like(g), np.zeros(n, np.int32), 1, 32, depth, 20, 0.0, 1e-6)
        msg = _compare_dict_to_oracle(mb.forest[0], rec.tree, "root")
        assert msg is None, msg
    else:
        assert np.array_equal(mb.predict_proba(X), mp.predict_proba(X))


def test_fused_split_is_deterministic_across_runs():
    """The fused split writes a child's rows in the order its chunks claim
    their places, which varies with scheduling; histograms and leaf sums are
    exact, so repeated fits give the same node tables, leaves and training
    scores bit for bit."""
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1(n=200000, f=40, seed=9)
    kw = dict(objective="regression", num_bins=32, max_depth=6, n_estimators=6, learning_rate=0.3)
    runs = []
    for _ in range(3):
        m = WarpGBM(engine="batched", **kw).fit(X, y)
        runs.append((m._last_tables, m.predict(X)))
    for t, p in runs[1:]:
        for key in ("status", "feat", "bin", "value", "count"):
            assert np.array_equal(t[key], runs[0][0][key]), key
        assert np.array_equal(p, runs[0][1])


@pytest.mark.parametrize("case", ["one_era", "early_leaves", "depth10", "eras", "deep", "eval_set"])
def test_training_scores_after_fit_equal_plain(case):
    """The scores a fit leaves behind (what a continued fit starts from) are
    the per-node engine's bit for bit, whichever way the batched engine adds
    its trees: from the leaves its partition records (one era; with leaves
    decided above the last level; at depth 10, the deepest whose levels all
    fuse), by walking the codes (eras; a depth whose deepest levels do not
    fuse), or at the end of each round (an eval set, whose rounds log the
    training loss)."""
    from warpgbm_amd import WarpGBM
    X, y = _data_cfg1(n=30000, f=40, seed=7)
    kw = dict(objective="regression", num_bins=32, max_depth=5, n_estimators=9, learning_rate=0.2, random_state=0)
    fk = {}
    if case == "eras":
        fk["era_id"] = np.arange(len(y)) % 3
    elif case == "early_leaves":
        kw.update(max_depth=6, min_child_weight=2500)  # nodes under 5000 rows stop splitting
    elif case == "depth10":
        kw.update(max_depth=10, min_child_weight=1, n_estimators=4)
    elif case == "deep":
        kw.update(max_depth=11, min_child_weight=1, n_estimators=3)
    elif case == "eval_set":
        fk.update(X_eval=X[:2000], y_eval=y[:2000], eval_every_n_trees=4)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, **fk)
    mp = WarpGBM(engine="plain", **kw).fit(X, y, **fk)
    for i in range(kw["n_estimators"]):
        msg = _trees_equal(mb.forest[i], mp.forest[i], f"tree{i}:")
        assert msg is None, msg
    if case == "early_leaves":  # the case exercises leaves recorded above the last split level
        def leaf_depths(t, d=0):
            return [d] if "leaf_value" in t else leaf_depths(t["left"], d + 1) + leaf_depths(t["right"], d + 1)
        assert min(min(leaf_depths(t)) for t in mb.forest) < kw["max_depth"] - 1
    sb, sp = mb.gradients.cpu().numpy(), mp.gradients.cpu().numpy()
    assert np.array_equal(sb, sp), float(np.abs(sb - sp).max())
    # and they are the model's own raw predictions of the training rows
    np.testing.assert_allclose(sb, mb.predict(X), rtol=1e-5, atol=1e-5)
