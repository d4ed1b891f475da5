"""Multi-output regression (shared splits, vector leaves) on the device.

The behaviour checks follow the reference's tests/test_multi_regression.py
(shapes and vector leaves, NaN targets vs explicit masking weights, 1-D vs
2-D weights, vector model vs one model per output, warm start, chunked
predict); the parity check grows the oracle's shared-split tree from the
engine's own float32 targets (teacher forcing) and compares node by node.
"""
import numpy as np
import pytest

from oracle import gbm_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
sk = pytest.importorskip("sklearn.datasets")


def _any_leaf(node):
    return node["leaf_value"] if "leaf_value" in node else _any_leaf(node["left"])


def _model(**kw):
    from warpgbm_amd import WarpGBM
    return WarpGBM(objective="multi_regression", device="cuda", **kw)


def test_fit_predict_shapes_and_vector_leaves():
    X, y = sk.make_regression(n_samples=600, n_features=14, n_targets=3, noise=0.3, random_state=42)
    X, y = X.astype(np.float32), y.astype(np.float32)
    m = _model(n_estimators=35, max_depth=4, learning_rate=0.1, num_bins=32, random_state=42).fit(X, y)
    p = m.predict(X)
    assert p.shape == y.shape and m.num_outputs == 3
    leaf = _any_leaf(m.forest[0])
    assert isinstance(leaf, list) and len(leaf) == 3
    assert np.mean((p - y) ** 2) < 0.5 * np.mean((y - y.mean(axis=0)) ** 2)


def test_nan_targets_match_explicit_masking_weights():
    X, y = sk.make_regression(n_samples=500, n_features=10, n_targets=3, noise=0.2, random_state=7)
    X, y = X.astype(np.float32), y.astype(np.float32)
    nan_mask = np.random.default_rng(7).random(y.shape) < 0.25
    y_nan = y.copy()
    y_nan[nan_mask] = np.nan
    kw = dict(n_estimators=30, max_depth=4, learning_rate=0.1, num_bins=32, min_child_weight=1, random_state=7)
    p_nan = _model(**kw).fit(X, y_nan).predict(X)
    p_mask = _model(**kw).fit(X, np.nan_to_num(y_nan, nan=0.0),
                             sample_weight=(~nan_mask).astype(np.float32)).predict(X)
    assert np.isfinite(p_nan).all() and np.isfinite(p_mask).all()
    obs = ~nan_mask
    ratio = np.mean((p_nan[obs] - y[obs]) ** 2) / max(np.mean((p_mask[obs] - y[obs]) ** 2), 1e-8)
    assert 0.8 < ratio < 1.25


def test_uniform_1d_and_2d_weights_train_the_same_model():
    X, y = sk.make_regression(n_samples=400, n_features=12, n_targets=2, noise=0.15, random_state=17)
    X, y = X.astype(np.float32), y.astype(np.float32)
    w1 = np.full(X.shape[0], 2.5, np.float32)
    kw = dict(n_estimators=25, max_depth=3, learning_rate=0.1, random_state=17)
    p1 = _model(**kw).fit(X, y, sample_weight=w1).predict(X)
    p2 = _model(**kw).fit(X, y, sample_weight=np.repeat(w1[:, None], 2, axis=1)).predict(X)
    np.testing.assert_array_equal(p1, p2)


def test_vector_model_close_to_one_model_per_output():
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(123)
    n, f = 1200, 16
    X = rng.normal(size=(n, f)).astype(np.float32)
    lat = 2.0 * X[:, 0] - 1.4 * X[:, 1] + 0.6 * X[:, 2]
    y = np.column_stack([lat + 0.15 * rng.normal(size=n), 0.8 * lat + 0.7 * X[:, 3] + 0.15 * rng.normal(size=n),
                         -0.5 * lat + 0.4 * X[:, 4] + 0.15 * rng.normal(size=n)]).astype(np.float32)
    tr, te = np.arange(840), np.arange(840, n)
    kw = dict(n_estimators=40, max_depth=4, learning_rate=0.1, num_bins=32, random_state=123, device="cuda")
    pm = WarpGBM(objective="multi_regression", **kw).fit(X[tr], y[tr]).predict(X[te])
    pi = np.column_stack([WarpGBM(objective="regression", **kw).fit(X[tr], y[tr, k]).predict(X[te])
                          for k in range(3)])
    ratio = np.mean((pm - y[te]) ** 2, axis=0) / np.maximum(np.mean((pi - y[te]) ** 2, axis=0), 1e-8)
    assert np.all(ratio < 1.6), ratio


def test_warm_start_equals_full_training_bit_for_bit():
    X, y = sk.make_regression(n_samples=500, n_features=12, n_targets=3, noise=0.2, random_state=31)
    X, y = X.astype(np.float32), y.astype(np.float32)
    kw = dict(max_depth=4, learning_rate=0.1, random_state=31)
    full = _model(n_estimators=50, **kw).fit(X, y)
    warm = _model(n_estimators=25, warm_start=True, **kw).fit(X, y)
    warm.n_estimators = 50
    warm.fit(X, y)
    np.testing.assert_array_equal(full.predict(X), warm.predict(X))


def test_chunked_predict_is_bit_identical_and_save_load_round_trip(tmp_path):
    from warpgbm_amd import WarpGBM
    X, y = sk.make_regression(n_samples=700, n_features=9, n_targets=2, noise=0.25, random_state=87)
    X, y = X.astype(np.float32), y.astype(np.float32)
    m = _model(n_estimators=30, max_depth=4, learning_rate=0.1, random_state=87).fit(X, y)
    full = m.predict(X)
    np.testing.assert_array_equal(full, np.vstack([m.predict(X[:350]), m.predict(X[350:])]))
    path = str(tmp_path / "multi.pkl")
    m.save_model(path)
    m2 = WarpGBM(objective="multi_regression", device="cuda").load_model(path)
    np.testing.assert_array_equal(m2.predict(X), full)


def test_eval_set_losses_and_early_stopping():
    X, y = sk.make_regression(n_samples=800, n_features=10, n_targets=2, noise=5.0, random_state=3)
    X, y = X.astype(np.float32), y.astype(np.float32)
    m = _model(n_estimators=60, max_depth=3, learning_rate=0.3, random_state=3)
    m.fit(X[:600], y[:600], X_eval=X[600:], y_eval=y[600:], eval_every_n_trees=1, early_stopping_rounds=3)
    assert len(m.training_loss) == m._trees_trained and len(m.eval_loss) == m._trees_trained
    p = m.predict(X[600:]).astype(np.float64)
    assert abs(m.eval_loss[-1] - np.mean((p - y[600:]) ** 2)) < 1e-4 * np.mean((p - y[600:]) ** 2)


@pytest.mark.parametrize("eras", [1, 4])
def test_trees_match_the_oracle_with_nan_targets_weights_and_eras(eras):
    """Teacher forcing: from the engine's float32 scores, rebuild the float32
    targets and their per-row sums as the gradient kernel forms them, grow
    the oracle's shared-split tree on them and compare every node: same
    (feature, bin), same leaf sizes, each leaf vector within 1e-5 relative;
    the replayed float32 scores equal the engine's predictions on the
    training matrix bit for bit, and the predictions lie within 1e-5 of the
    oracle's float64 forest evaluation."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(5 + eras)
    N, F, M, B, D, T = 6000, 12, 3, 32, 4, 8
    X = rng.standard_normal((N, F)).astype(np.float32)
    Y = np.stack([X[:, m] * (m + 1) + np.sin(X[:, m + 3]) + 0.1 * rng.standard_normal(N) for m in range(M)],
                 axis=1).astype(np.float32)
    Y[rng.random(Y.shape) < 0.1] = np.nan
    W = rng.uniform(0.5, 2.0, (N, M)).astype(np.float32)
    era = rng.integers(0, eras, N).astype(np.int32)
    lr = np.float32(0.2)
    m = WarpGBM(objective="multi_regression", num_bins=B, max_depth=D, n_estimators=T, learning_rate=0.2,
                min_child_weight=20, random_state=0, device="cuda").fit(X, Y, era_id=era, sample_weight=W)
    Wn = (W / np.float32(W.mean())).astype(np.float32)  # mean-normalised weights
    codes, _ = O.bin_matrix(X, B)
    einv = np.unique(era, return_inverse=True)[1]
    valid = np.isfinite(Y)
    scores = np.tile(m.base_prediction.cpu().numpy().astype(np.float32), (N, 1))
    for t in range(T):
        g = np.where(valid, (np.where(valid, Y, 0) - scores) * Wn, 0).astype(np.float32)
        h = np.where(valid, Wn, 0).astype(np.float32)
        sg, sh = np.zeros(N, np.float32), np.zeros(N, np.float32)
        for k in range(M):
            sg = (sg + g[:, k]).astype(np.float32)
            sh = (sh + h[:, k]).astype(np.float32)
        rec = O.grow_tree_multi(codes, g, h, einv, eras, B, D, 20, 0.0, 1e-6, split_g=sg.astype(np.float64),
                                split_h=sh.astype(np.float64))
        _compare(m.forest[t], rec.tree, f"tree{t}:")
        for heap, node in _leaf_rows(m.forest[t], codes).items():
            rows, vec = node
            scores[rows] = (scores[rows] + (lr * np.asarray(vec, np.float32)).astype(np.float32)).astype(np.float32)
    pred = m.predict(X)
    np.testing.assert_array_equal(pred, scores)
    trees = [_plain(m.forest[t]) for t in range(T)]
    ref = O.predict_multi(codes, trees, 0.2, m.base_prediction.cpu().numpy())
    np.testing.assert_allclose(pred, ref, rtol=1e-5, atol=1e-5)


def _plain(t):
    if "leaf_value" in t:
        return dict(t)
    return {"feature": int(t["feature"].item()), "bin": t["bin"], "left": _plain(t["left"]),
            "right": _plain(t["right"])}


def _leaf_rows(tree, codes):
    out = {}

    def walk(node, rows, heap):
        if "leaf_value" in node:
            out[heap] = (rows, node["leaf_value"])
            return
        f = int(node["feature"].item()) if torch.is_tensor(node["feature"]) else int(node["feature"])
        sel = codes[rows, f] <= node["bin"]
        walk(node["left"], rows[sel], 2 * heap + 1)
        walk(node["right"], rows[~sel], 2 * heap + 2)

    walk(tree, np.arange(codes.shape[0]), 0)
    return out


def _compare(t, o, path):
    if "leaf_value" in o:
        assert "leaf_value" in t, f"{path}: engine split where the oracle has a leaf"
        assert t["samples"] == o["samples"], f"{path}: {t['samples']} rows vs {o['samples']}"
        got, ref = np.asarray(t["leaf_value"]), np.asarray(o["leaf_value"])
        assert got.shape == ref.shape, f"{path}: leaf vector {got.shape} vs {ref.shape}"
        assert np.all(np.abs(got - ref) <= 1e-5 * np.maximum(np.abs(ref), 1e-3)), f"{path}: {got} vs {ref}"
        return
    assert "feature" in t, f"{path}: engine leaf where the oracle splits"
    f = int(t["feature"].item())
    assert (f, t["bin"]) == (o["feature"], o["bin"]), f"{path}: ({f},{t['bin']}) vs ({o['feature']},{o['bin']})"
    _compare(t["left"], o["left"], path + "L")
    _compare(t["right"], o["right"], path + "R")
