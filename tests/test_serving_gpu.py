"""GPU tests of the compiled heap forest and the continuous-batching server.

* the heap forest predicts what the reference op predicts on the flattened
  forest (within 1e-5) and what the FP64 oracle predicts;
* a request's result is bit-identical served alone or mixed with others,
  under concurrent submission (continuous batching);
* admission respects the memory budget (oversized requests are rejected).
"""
import threading

import numpy as np
import pytest

from oracle import gbm_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def model_and_codes():
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(11)
    n, f = 20000, 40
    X = rng.standard_normal((n, f)).astype(np.float32)
    y = (np.tanh(X[:, :8] @ rng.standard_normal(8)) + 0.1 * rng.standard_normal(n)).astype(np.float32)
    m = WarpGBM(objective="regression", num_bins=32, max_depth=6, n_estimators=30, learning_rate=0.1).fit(X, y)
    codes = m.bin_inference_data(X)
    return m, codes


def test_heap_forest_matches_reference_op_and_oracle(model_and_codes):
    from warpgbm_amd.serving import HeapForest
    m, codes = model_and_codes
    hf = HeapForest.from_model(m, tree_block=7)
    out = hf.predict(codes)[:, 0]
    ref = m.predict_binned(codes)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), ref.cpu().numpy(), rtol=1e-5, atol=1e-5)
    c = codes.cpu().numpy()
    flats = [O.flatten_tree(_plain(t), 2 ** 7) for t in m.forest]
    oracle = O.predict_forest_flat(c, flats, 0.1, base=m.base_prediction)
    np.testing.assert_allclose(out.cpu().numpy(), oracle, rtol=1e-5, atol=1e-5)


def _plain(t):
    if "leaf_value" in t:
        return dict(t)
    return {"feature": int(t["feature"].item()), "bin": t["bin"], "left": _plain(t["left"]),
            "right": _plain(t["right"])}


def test_multiclass_heap_forest_matches_predict_proba():
    from warpgbm_amd import WarpGBM
    from warpgbm_amd.serving import HeapForest
    rng = np.random.default_rng(3)
    n, f, K = 5000, 16, 3
    X = rng.standard_normal((n, f)).astype(np.float32)
    lab = np.argmax(X[:, :6] @ rng.standard_normal((6, K)) + rng.gumbel(size=(n, K)), axis=1)
    m = WarpGBM(objective="multiclass", num_bins=16, max_depth=4, n_estimators=10).fit(X, lab)
    hf = HeapForest.from_model(m)
    codes = m.bin_inference_data(X)
    logits = hf.predict(codes)
    p = torch.softmax(logits, dim=1).cpu().numpy()
    np.testing.assert_allclose(p, m.predict_proba(X), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("piece_rows", [None, 1024])
def test_server_alone_equals_mixed(model_and_codes, piece_rows):
    """Requests served one at a time equal the same requests sent by four
    concurrent clients (launches shared), whole or cut into 1024-row
    pieces spread over launches."""
    from warpgbm_amd.serving import HeapForest
    from warpgbm_amd.server import PredictServer
    m, codes = model_and_codes
    hf = HeapForest.from_model(m)
    c = codes.cpu().numpy()
    rng = np.random.default_rng(0)
    sizes = rng.integers(1, 3000, size=40)
    starts = rng.integers(0, c.shape[0] - 3000, size=40)
    reqs = [c[s:s + k] for s, k in zip(starts, sizes)]
    alone = []
    srv = PredictServer(hf, memory_budget=64 << 20, max_batch_rows=8192, slots=3, piece_rows=piece_rows)
    try:
        for r in reqs:  # one at a time
            alone.append(srv.submit(r).result(timeout=60))
        futs = [None] * len(reqs)

        def client(ids):
            for i in ids:
                futs[i] = srv.submit(reqs[i])

        ths = [threading.Thread(target=client, args=(list(range(k, 40, 4)),)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        mixed = [f.result(timeout=60) for f in futs]
    finally:
        srv.close()
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)
    assert srv.batches < 2 * len(reqs)  # the concurrent phase shared launches
    ref = m.predict_binned(codes).cpu().numpy()
    for r, s, k in zip(alone, starts, sizes):
        np.testing.assert_allclose(r[:, 0], ref[s:s + k], rtol=1e-5, atol=1e-5)


def test_server_rejects_over_budget(model_and_codes):
    from warpgbm_amd.serving import HeapForest
    from warpgbm_amd.server import PredictServer
    m, codes = model_and_codes
    srv = PredictServer(HeapForest.from_model(m), memory_budget=100_000, max_batch_rows=4096, slots=1)
    try:
        with pytest.raises(ValueError, match="budget"):
            srv.submit(codes[:4000].cpu().numpy())
        small = srv.submit(codes[:10].cpu().numpy()).result(timeout=30)
        assert small.shape == (10, 1)
    finally:
        srv.close()


@pytest.mark.parametrize("tree_block,depth_trees", [(64, 30), (7, 30), (16, 30)])
def test_traversal_variants_equal_bit_for_bit(model_and_codes, tree_block, depth_trees):
    """Node reads through the cache (2 or 4 trees in flight per lane) and
    nodes staged in LDS add every row's trees in the same order."""
    from warpgbm_amd import _lib
    from warpgbm_amd.serving import HeapForest
    m, codes = model_and_codes
    hf = HeapForest.from_model(m, tree_block=tree_block)
    outs = []
    try:
        for mode in (1, 2, 3):
            _lib.call("wgbm_set_predict_mode", mode)
            outs.append(hf.predict(codes[:1000 + 37])[:, 0].cpu().numpy())
    finally:
        _lib.call("wgbm_set_predict_mode", 0)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_tree_order_heap_predict_equals_reference_op_bitwise(model_and_codes):
    """The estimator's predict path (compiled heap forest, trees added in
    tree order) equals the reference op on the flattened forest bit for bit,
    for the trained model and for the same model saved and loaded (compiled
    from its tree dicts)."""
    import os
    import tempfile
    from warpgbm_amd import WarpGBM
    from warpgbm_amd.cuda import node_kernel as nk
    from warpgbm_amd.serving import HeapForest
    m, codes = model_and_codes
    n = codes.shape[0]
    tt = m._get_cached_regression_tree_tensor()
    ref = torch.full((n,), float(np.float32(m.base_prediction)), dtype=torch.float32, device=codes.device)
    nk.predict_forest(codes, tt.contiguous(), m.learning_rate, ref)
    hf = HeapForest.from_model(m)
    assert torch.equal(hf.predict_tree_order(codes)[:, 0], ref)
    assert torch.equal(m.predict_binned(codes), ref)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.pkl")
        m.save_model(p)
        m2 = WarpGBM().load_model(p)
        m2.device = m.device
        assert torch.equal(m2.predict_binned(codes), ref)
