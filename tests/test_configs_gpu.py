"""GPU parity at the benchmark workloads' own sizes.

* era splitting, 500k x 400 pre-binned int8 codes in [0, 5), 200 lognormal
  eras, depth 8: batched trees equal plain ones; tree 0's splits through
  depth 4 and their per-era gains equal the FP64 oracle's (DES rule,
  warpgbm/core.py:1322-1328);
* multiclass, 200k x 100, 5 classes, depth 6: all class trees of rounds 0-2
  equal the oracle's, grown from the engine's own float32 softmax targets
  (teacher forcing, core.py:1719-1791); predict_proba within 1e-5 of the
  oracle's;
* serving, a 1000-tree depth-8 forest over 400 features: the heap forest
  predicts 10k rows within 1e-5 of the oracle's flattened-forest traversal
  (warpgbm/cuda/predict.cu:5-45), and the server returns every request of a
  log-uniform 1..100k mix bit for bit as it does alone.
"""
import threading

import numpy as np
import pytest

from oracle import gbm_oracle as O
import _parity as PAR

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _plain(t):
    if "leaf_value" in t:
        return dict(t)
    f = t["feature"]
    return {"feature": int(f.item()) if torch.is_tensor(f) else int(f), "bin": int(t["bin"]),
            "left": _plain(t["left"]), "right": _plain(t["right"])}


def _same_tree(a, b):
    a, b = _plain(a), _plain(b)
    return a == b


def _compare_to_depth(t, o, max_d, path="root", d=0, out=None):
    """Engine tree t against oracle tree o on every node of depth < max_d:
    same split, or both leaves with the same rows and values within 1e-5."""
    if out is None:
        out = []
    if d >= max_d:
        return out
    if "leaf_value" in o:
        assert "leaf_value" in t, f"{path}: engine split where oracle has a leaf"
        assert t["samples"] == o["samples"], path
        assert abs(t["leaf_value"] - o["leaf_value"]) <= 1e-5 * max(abs(o["leaf_value"]), 1e-3), path
        return out
    assert "feature" in t, f"{path}: engine leaf where oracle splits"
    assert (t["feature"], t["bin"]) == (o["feature"], o["bin"]), \
        f"{path}: split ({t['feature']},{t['bin']}) vs ({o['feature']},{o['bin']})"
    out.append(path)
    _compare_to_depth(t["left"], o["left"], max_d, path + "L", d + 1, out)
    _compare_to_depth(t["right"], o["right"], max_d, path + "R", d + 1, out)
    return out


def test_era_workload_500k_x_400_200_eras():
    from warpgbm_amd import WarpGBM
    from warpgbm_amd.synthetic import era_workload
    X, y, era = era_workload()
    E = 200
    kw = dict(objective="regression", num_bins=5, max_depth=8, n_estimators=2, learning_rate=0.01)
    mb = WarpGBM(engine="batched", **kw).fit(X, y, era_id=era)
    tables = mb._last_tables
    mp = WarpGBM(engine="plain", **kw).fit(X, y, era_id=era)
    for i in range(2):
        assert _same_tree(mb.forest[i], mp.forest[i]), f"tree {i}"
    np.testing.assert_array_equal(mb.per_era_feature_importance_, mp.per_era_feature_importance_)
    # tree 0 against the oracle grown from the same float32 targets, splits
    # at depths 0..4 (the oracle stops at depth 5)
    g = (y - np.float32(mb.base_prediction)).astype(np.float32)
    rec = O.grow_tree(X, g, np.ones_like(g), era, E, 5, 5, 20, 0.0, 1e-6)
    splits = _compare_to_depth(_plain(mb.forest[0]), rec.tree, 5)
    assert len(splits) >= 20
    # per-era gains of those splits (the DES vote's inputs), heap order
    eg = tables["era_gain"][0, 0]  # [heap slots, E], round 0
    worst = 0.0
    for heap, node in rec.nodes.items():
        if node[0] != "split":
            continue
        ref = node[3]
        err = np.abs(eg[heap].astype(np.float64) - ref).max() / max(np.abs(ref).max(), 1e-12)
        worst = max(worst, err)
    PAR.record("era workload per-era split gains (relative to the node's largest)", worst, 1e-5)
    assert worst <= 1e-5


def test_multiclass_workload_200k_x_100_k5_rounds_0_to_2():
    from warpgbm_amd import WarpGBM, _lib
    from warpgbm_amd.synthetic import multiclass_workload
    X, lab = multiclass_workload()
    n, K, R, D = X.shape[0], 5, 3, 6
    m = WarpGBM(objective="multiclass", num_bins=32, max_depth=D, n_estimators=R, learning_rate=0.1).fit(X, lab)
    codes, _ = O.bin_matrix(X, 32)
    assert np.array_equal(m.bin_inference_data(X).cpu().numpy(), codes)  # codes bit-exact at this size
    counts = np.bincount(lab, minlength=K)
    sc = np.empty((n, K), np.float32)
    for k in range(K):
        sc[:, k] = np.float32(np.log(np.float32(np.float32(counts[k]) / np.float32(n)) + np.float32(1e-10)))
    dev = torch.device("cuda")
    l_d = torch.from_numpy(lab.astype(np.int32)).to(dev)
    w_d = torch.ones(n, dtype=torch.float32, device=dev)
    g_d = torch.empty(K * n, dtype=torch.float32, device=dev)
    h_d = torch.empty(K * n, dtype=torch.float32, device=dev)
    eras = np.zeros(n, np.int32)
    lr = np.float32(0.1)
    recs = []
    for r in range(R):
        s_d = torch.from_numpy(sc).to(dev)
        _lib.call("wgbm_grad_softmax", s_d.data_ptr(), l_d.data_ptr(), w_d.data_ptr(), n, K, g_d.data_ptr(),
                  h_d.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        g = g_d.cpu().numpy().reshape(K, n)
        h = h_d.cpu().numpy().reshape(K, n)
        for k in range(K):
            t = _plain(m.forest[r][k])
            rec = O.grow_tree(codes, g[k], h[k], eras, 1, 32, D, 20, 0.0, 1e-6)
            recs.append(rec)
            _compare_to_depth(t, rec.tree, D + 1, f"round{r} class{k}")
            vals = O.predict_tree(t, codes).astype(np.float32)
            sc[:, k] = (sc[:, k] + (lr * vals).astype(np.float32)).astype(np.float32)
    # predict_proba against the oracle's (uniform log(1/K) prior, softmax)
    p = m.predict_proba(X)
    po = O.predict_proba_binned(codes, [[_plain(t) for t in rd] for rd in m.forest], 0.1, K)
    err = float(np.max(np.abs(p - po) / np.maximum(po, 1e-6)))
    PAR.record("multiclass predict_proba (relative)", err, 1e-5)
    assert err <= 1e-5
    # per-era importance (one era: the total gain) against the oracle's records
    imp = O.per_era_importance(recs, 1, X.shape[1])
    scale = np.maximum(np.abs(imp), 1e-12)
    ierr = float(np.max(np.abs(m.per_era_feature_importance_ - imp) / scale))
    PAR.record("multiclass per-era importance (relative)", ierr, 1e-5)
    assert ierr <= 1e-5


@pytest.fixture(scope="module")
def forest_1000x8():
    from warpgbm_amd import WarpGBM
    from warpgbm_amd.serving import HeapForest
    from warpgbm_amd.synthetic import regression
    X, y = regression(200_000, 400, seed=1)
    m = WarpGBM(objective="regression", num_bins=32, max_depth=8, n_estimators=1000, learning_rate=0.01).fit(X, y)
    tb = m._last_tables
    hf = HeapForest.from_tables(*(tb[k].reshape(-1, tb[k].shape[-1]) for k in ("status", "feat", "bin", "value")),
                                8, 0.01, 400, [m.base_prediction], device="cuda")
    codes = m.bin_inference_data(X)
    return m, hf, codes


def test_serving_forest_1000_trees_depth_8_matches_oracle(forest_1000x8):
    m, hf, codes = forest_1000x8
    assert hf.n_trees == 1000 and hf.depth == 8
    c = codes[:10000].contiguous()
    out = hf.predict(c)[:, 0].cpu().numpy()
    flats = [O.flatten_tree(_plain(t), 2 ** 9) for t in m.forest]
    ref = O.predict_forest_flat(c.cpu().numpy(), flats, 0.01, base=m.base_prediction)
    # A prediction is a float32 sum of 1001 terms (base + one leaf per
    # tree): its error is judged against the sum of the terms' magnitudes.
    for f in flats:
        f[:, 5] = np.abs(f[:, 5])
    scale = O.predict_forest_flat(c.cpu().numpy(), flats, 0.01, base=abs(m.base_prediction))
    err = float(np.max(np.abs(out - ref) / scale))
    PAR.record("serving heap forest 1000x8, 10k rows (relative to sum of |terms|)", err, 1e-5)
    assert err <= 1e-5
    # the reference op (trees added one by one) sums in another order
    op = m.predict_binned(c).cpu().numpy()
    assert float(np.max(np.abs(out - op) / scale)) <= 1e-5


def test_server_loguniform_mix_alone_equals_mixed(forest_1000x8):
    from warpgbm_amd.server import PredictServer
    m, hf, codes = forest_1000x8
    host = codes.cpu().numpy()
    rng = np.random.default_rng(5)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(1e5), 24))).astype(int).clip(1, 100_000)
    starts = rng.integers(0, host.shape[0] - sizes + 1)
    reqs = [host[s:s + k] for s, k in zip(starts, sizes)]
    srv = PredictServer(hf, memory_budget=1 << 30)
    try:
        alone = [srv.submit(r).result(timeout=60) for r in reqs]
        futs = [None] * len(reqs)

        def client(ids):
            for i in ids:
                futs[i] = srv.submit(reqs[i])

        ths = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        mixed = [f.result(timeout=60) for f in futs]
    finally:
        srv.close()
    for i, (a, b) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a, b), f"request {i} ({sizes[i]} rows)"
    one = hf.predict(codes[: int(sizes[0])].contiguous()).cpu().numpy() if starts[0] == 0 else None
    for i in range(3):  # and equal to one launch of that request alone
        single = hf.predict(torch.from_numpy(reqs[i]).to("cuda")).cpu().numpy()
        assert np.array_equal(alone[i], single)
    del one
