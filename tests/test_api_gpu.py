"""Behavioural acceptance tests of the drop-in estimator on the GPU.

These restate, in fresh code, the end-to-end properties the reference's own
test-suite asserts (warpgbm tests/: predict chunking, warm start, save/load,
monotone constraints, sample weights, random state, multiclass, importances,
fit quality, evaluation / early stopping).  This engine is deterministic, so
where the reference allows a tolerance for float-atomic noise these tests
ask for equality.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _reg(n=6000, f=16, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f)).astype(np.float32)
    y = (X[:, 0] * 2 + np.sin(X[:, 1]) + 0.5 * X[:, 2] * X[:, 3] + 0.1 * rng.standard_normal(n)).astype(np.float32)
    return X, y


def _cls(n=6000, f=12, K=3, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f)).astype(np.float32)
    lab = np.argmax(X[:, :6] @ rng.standard_normal((6, K)) + 0.3 * rng.gumbel(size=(n, K)), axis=1)
    return X, lab


@pytest.mark.parametrize("engine", ["batched", "plain"])
def test_predict_chunking(engine):
    """tests/test_predict_chunking.py:19-40: predicting in pieces equals one call."""
    from warpgbm_amd import WarpGBM
    X, y = _reg()
    m = WarpGBM(num_bins=32, max_depth=4, n_estimators=15, engine=engine).fit(X, y)
    full = m.predict(X)
    parts = np.concatenate([m.predict(X[:2500]), m.predict(X[2500:])])
    assert np.array_equal(full, parts)
    Xc, lab = _cls()
    mc = WarpGBM(objective="multiclass", num_bins=16, max_depth=3, n_estimators=8, engine=engine).fit(Xc, lab)
    pf = mc.predict_proba(Xc)
    pp = np.concatenate([mc.predict_proba(Xc[:3000]), mc.predict_proba(Xc[3000:])])
    assert np.array_equal(pf, pp)


@pytest.mark.parametrize("objective", ["regression", "multiclass"])
def test_warm_start_equals_one_fit(objective):
    """tests/test_warm_start.py:15-196: 6 + 6 trees with warm start == 12 trees."""
    from warpgbm_amd import WarpGBM
    X, y = _reg() if objective == "regression" else _cls()
    kw = dict(objective=objective, num_bins=16, max_depth=3, learning_rate=0.2)
    full = WarpGBM(n_estimators=12, **kw).fit(X, y)
    ws = WarpGBM(n_estimators=6, warm_start=True, **kw).fit(X, y)
    ws.n_estimators = 12
    ws.fit(X, y)
    assert ws._trees_trained == 12
    if objective == "regression":
        assert np.array_equal(ws.predict(X), full.predict(X))
    else:
        assert np.array_equal(ws.predict_proba(X), full.predict_proba(X))
    assert np.array_equal(ws.feature_importance_, full.feature_importance_)


def test_save_load_predicts_identically(tmp_path):
    """tests/test_warm_start.py:202-300: a saved and reloaded model predicts the same."""
    from warpgbm_amd import WarpGBM
    X, y = _reg()
    m = WarpGBM(num_bins=32, max_depth=4, n_estimators=10).fit(X, y)
    path = tmp_path / "m.pkl"
    m.save_model(str(path))
    m2 = WarpGBM().load_model(str(path))
    assert np.array_equal(m2.predict(X), m.predict(X))
    Xc, lab = _cls()
    mc = WarpGBM(objective="multiclass", num_bins=16, max_depth=3, n_estimators=5).fit(Xc, lab)
    mc.save_model(str(path))
    mc2 = WarpGBM().load_model(str(path))
    assert np.array_equal(mc2.predict_proba(Xc), mc.predict_proba(Xc))
    assert np.array_equal(mc2.predict(Xc), mc.predict(Xc))


def _mono_data(seed, n=600):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-2.0, 2.0, n).astype(np.float32)
    X = np.column_stack([x0, rng.normal(0.0, 0.3, n)]).astype(np.float32)
    return X, x0, rng


@pytest.mark.parametrize("case", ["increasing", "decreasing", "eras"])
def test_monotone_constraint_holds_on_a_grid(case):
    """tests/test_monotonic_constraints.py:27-120: along the constrained
    feature (the signal's feature) the predictions are monotone on a grid,
    without and with eras (two eras with different slopes); the list and the
    dict forms of the constraint give the same model."""
    from warpgbm_amd import WarpGBM
    X, x0, rng = _mono_data({"increasing": 7, "decreasing": 9, "eras": 33}[case])
    noise = 0.04 * rng.standard_normal(len(x0)).astype(np.float32)
    era = None
    if case == "increasing":
        y, sign = 3.0 * x0 + noise, 1
    elif case == "decreasing":
        y, sign = -3.0 * x0 + noise, -1
    else:
        era = (x0 > 0).astype(np.int32)
        y, sign = np.where(era == 0, 2.0 * x0, 0.8 * x0) + noise, 1
    y = y.astype(np.float32)
    kw = dict(num_bins=32, max_depth=3, n_estimators=40, learning_rate=0.1, random_state=1)
    m = WarpGBM(monotonic_constraints=[sign, 0], **kw).fit(X, y, era_id=era)
    md = WarpGBM(monotonic_constraints={0: sign}, **kw).fit(X, y, era_id=era)
    grid = np.linspace(-2.0, 2.0, 240).astype(np.float32)
    Q = np.column_stack([grid, np.zeros_like(grid)]).astype(np.float32)
    p = m.predict(Q)
    assert np.all(sign * np.diff(p) >= -1e-5)
    assert np.array_equal(p, md.predict(Q))
    assert np.array_equal(m.predict(X), md.predict(X))


def test_uniform_weights_equal_unweighted():
    """tests/test_sample_weight.py:7-76: a constant weight changes nothing."""
    from warpgbm_amd import WarpGBM
    X, y = _reg()
    kw = dict(num_bins=32, max_depth=4, n_estimators=10)
    a = WarpGBM(**kw).fit(X, y)
    b = WarpGBM(**kw).fit(X, y, sample_weight=np.full(len(y), 3.5, np.float32))
    assert np.array_equal(a.predict(X), b.predict(X))


def test_weighted_mean_tracked_without_splits():
    """tests/test_sample_weight.py:142-178: with min_child_weight above the
    total weight nothing splits, and the model converges to the weighted mean."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(4)
    n = 2000
    X = rng.standard_normal((n, 5)).astype(np.float32)
    y = np.where(rng.random(n) < 0.5, 0.0, 10.0).astype(np.float32)
    w = np.where(y > 5, 3.0, 1.0).astype(np.float32)
    m = WarpGBM(num_bins=16, max_depth=3, n_estimators=60, learning_rate=0.3, min_child_weight=n * 10).fit(
        X, y, sample_weight=w)
    assert all("leaf_value" in t for t in m.forest)
    target = float(np.sum(w * y) / np.sum(w))
    assert abs(float(np.mean(m.predict(X))) - target) < 0.15


def test_random_state_reproducible_with_column_sampling():
    """tests/test_random_state.py:7-135: same seed, same model."""
    from warpgbm_amd import WarpGBM
    X, y = _reg()
    kw = dict(num_bins=32, max_depth=4, n_estimators=10, colsample_bytree=0.5, random_state=7)
    a = WarpGBM(**kw).fit(X, y)
    b = WarpGBM(**kw).fit(X, y)
    assert np.array_equal(a.predict(X), b.predict(X))
    c = WarpGBM(**dict(kw, random_state=8)).fit(X, y)
    assert not np.array_equal(a.predict(X), c.predict(X))


def test_multiclass_string_labels_probabilities_accuracy():
    """tests/test_multiclass.py: string labels round-trip, rows sum to 1, the
    model learns a separable problem."""
    from warpgbm_amd import WarpGBM
    X, lab = _cls(K=4)
    names = np.array(["alpha", "beta", "gamma", "delta"])[lab]
    m = WarpGBM(objective="multiclass", num_bins=32, max_depth=4, n_estimators=30, learning_rate=0.2).fit(X, names)
    p = m.predict_proba(X)
    np.testing.assert_allclose(p.sum(axis=1), 1.0, atol=1e-5)
    pred = m.predict(X)
    assert set(np.unique(pred)) <= set(names)
    assert np.mean(pred == names) > 0.8


def test_single_era_importance_equals_total():
    """tests/test_feature_importance.py:265-298: one era -> per-era == total."""
    from warpgbm_amd import WarpGBM
    X, y = _reg()
    m = WarpGBM(num_bins=32, max_depth=4, n_estimators=10).fit(X, y)
    per = m.get_per_era_feature_importance(normalize=False)
    assert per.shape == (1, X.shape[1])
    np.testing.assert_allclose(per[0], m.get_feature_importance(normalize=False))
    assert np.argmax(m.get_feature_importance()) == 0


def test_fit_quality_correlation():
    """tests/test_fit_predict_corr.py:8-52 (smaller): held-out correlation > 0.9."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(9)
    n, f = 60000, 100
    X = rng.standard_normal((n, f)).astype(np.float32)
    y = (X[:, :10] @ rng.standard_normal(10) + 0.3 * rng.standard_normal(n)).astype(np.float32)
    m = WarpGBM(num_bins=32, max_depth=5, n_estimators=200, learning_rate=0.1).fit(X[:50000], y[:50000])
    assert np.corrcoef(m.predict(X[50000:]), y[50000:])[0, 1] > 0.9


@pytest.mark.parametrize("objective", ["regression", "multiclass"])
def test_eval_losses_match_plain(objective):
    """The batched engine keeps eval scores on the device (each call adds its
    new trees) and computes every tree's training loss on the device, running
    the rounds between evaluations in one call; the losses it logs match the
    plain engine's torch-computed ones (core.py:1488-1516, :1793-1845)."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(21)
    n, f = 6000, 20
    X = rng.standard_normal((n, f)).astype(np.float32)
    Xe = rng.standard_normal((1500, f)).astype(np.float32)
    if objective == "regression":
        w = rng.standard_normal(5)
        y = (X[:, :5] @ w + 0.2 * rng.standard_normal(n)).astype(np.float32)
        ye = (Xe[:, :5] @ w).astype(np.float32)
        kw = dict(objective="regression", learning_rate=0.2)
    else:
        W = rng.standard_normal((5, 3))
        y = np.argmax(X[:, :5] @ W + rng.gumbel(size=(n, 3)), axis=1)
        ye = np.argmax(Xe[:, :5] @ W, axis=1)
        kw = dict(objective="multiclass", learning_rate=0.3)
    fits = {}
    for engine in ("batched", "plain"):
        m = WarpGBM(num_bins=16, max_depth=3, n_estimators=23, engine=engine, **kw)
        m.fit(X, y, X_eval=Xe, y_eval=ye, eval_every_n_trees=4)
        fits[engine] = m
    b, p = fits["batched"], fits["plain"]
    assert b._trees_trained == p._trees_trained == 23
    assert len(b.training_loss) == len(p.training_loss) == 23
    assert len(b.eval_loss) == len(p.eval_loss) == 6  # trees 0, 4, 8, 12, 16, 20
    np.testing.assert_allclose(b.training_loss, p.training_loss, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(b.eval_loss, p.eval_loss, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("engine", ["batched", "plain"])
def test_eval_set_and_early_stopping(engine):
    """core.py:1488-1516: eval loss recorded every n trees; early stopping
    ends training once the eval loss rises; both engines stop at the same tree."""
    from warpgbm_amd import WarpGBM
    X, y = _reg(n=4000)
    rng = np.random.default_rng(11)
    Xe = rng.standard_normal((1000, X.shape[1])).astype(np.float32)
    ye = rng.standard_normal(1000).astype(np.float32)  # unrelated: eval loss soon rises
    m = WarpGBM(num_bins=32, max_depth=4, n_estimators=200, learning_rate=0.3, engine=engine)
    m.fit(X, y, X_eval=Xe, y_eval=ye, eval_every_n_trees=1, early_stopping_rounds=3)
    assert 0 < m._trees_trained < 200
    assert len(m.eval_loss) == m._trees_trained
    ref = WarpGBM(num_bins=32, max_depth=4, n_estimators=200, learning_rate=0.3, engine="plain")
    ref.fit(X, y, X_eval=Xe, y_eval=ye, eval_every_n_trees=1, early_stopping_rounds=3)
    assert ref._trees_trained == m._trees_trained


def test_integer_codes_of_any_width_and_empty_integer_input():
    """Pre-binned codes given as int32 train and predict exactly like the
    same int8 codes (the pre-binned bypass keys on integer dtype and
    max < num_bins); zero rows of integer input raise ValueError at the
    pre-binned check, as the reference's does (warpgbm/core.py:1967-1969)."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(71)
    X8 = rng.integers(0, 6, (4000, 9)).astype(np.int8)
    y = (X8[:, 0] - 0.5 * X8[:, 4] + 0.3 * rng.standard_normal(4000)).astype(np.float32)
    kw = dict(objective="regression", num_bins=6, max_depth=3, n_estimators=3, learning_rate=0.2)
    m8 = WarpGBM(**kw).fit(X8, y)
    m32 = WarpGBM(**kw).fit(X8.astype(np.int32), y)
    np.testing.assert_array_equal(m8.predict(X8), m32.predict(X8.astype(np.int32)))
    with pytest.raises(ValueError):
        m8.predict(X8[:0])


def test_staged_host_copies_equal_direct_copies(monkeypatch):
    """Host arrays copied through the pinned staging buffers (here with 4 KB
    chunks, so ragged last chunks and many buffer reuses occur) equal direct
    copies for float32, float64 -> float32, int32 and int8 data."""
    import warpgbm_amd.core as core
    monkeypatch.setattr(core, "_STAGE_BYTES", 4096)
    dev = torch.device("cuda")
    rng = np.random.default_rng(5)
    cases = [(rng.standard_normal((3001, 7)).astype(np.float32), torch.float32),
             (rng.standard_normal((2999, 5)), torch.float32),
             (rng.integers(-9, 9, (4097, 3)).astype(np.int32), None),
             (rng.integers(0, 5, (12345, 2)).astype(np.int8), None)]
    for a, dt in cases:
        got = core._to_device(a, dev, dt).cpu()
        ref = torch.from_numpy(a).to(dt if dt is not None else torch.from_numpy(a).dtype)
        assert got.dtype == ref.dtype and got.shape == ref.shape
        assert torch.equal(got, ref)


def test_autotune_installs_a_config_and_trees_do_not_change(tmp_path, monkeypatch):
    """autotune.tune times every candidate histogram configuration for a
    shape, installs the fastest and caches it; the trees grown under it are
    the trees of the automatic configuration (exact histograms), and a
    second call is served from the cache."""
    from warpgbm_amd import WarpGBM, autotune
    monkeypatch.setenv("WARPGBM_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    X, y = _reg(n=6000, f=16, seed=9)
    kw = dict(objective="regression", num_bins=32, max_depth=4, n_estimators=3, learning_rate=0.1)
    ref = WarpGBM(**kw).fit(X, y)
    try:
        best, times = autotune.tune(16, 32, rows=20000, use_cache=True)
        assert best in times and len(times) >= 2
        tuned = WarpGBM(**kw).fit(X, y)
        assert np.array_equal(ref.predict(X), tuned.predict(X))
        again, times2 = autotune.tune(16, 32, rows=20000, use_cache=True)
        assert again == best and times2 == {}
        assert (tmp_path / "tune.json").exists()
    finally:
        autotune.reset()


def test_fit_tunes_the_histogram_tiling_per_shape(tmp_path, monkeypatch):
    """A batched fit times the candidate tilings on a few rounds of its own
    data (one LDS plane with equal weights, two with weights), installs the
    fastest for that histogram shape only, and grows the trees an untuned
    fit grows; the next fit of the shape takes the choice from the cache.
    A tiling chosen for one shape never reaches another: after tuning at 5
    bins, a 127-bin fit still runs."""
    from warpgbm_amd import WarpGBM, autotune, _lib
    monkeypatch.setenv("WARPGBM_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    X, y = _reg(n=20000, f=64, seed=4)
    kw = dict(objective="regression", num_bins=32, max_depth=5, n_estimators=4, learning_rate=0.1)
    try:
        ref = WarpGBM(autotune=False, **kw).fit(X, y)
        tuned = WarpGBM(**kw).fit(X, y)
        info = tuned.hist_config_
        assert info is not None and not info["cached"] and info["candidates"] >= 4
        assert tuple(autotune.current(64, 32, 64, True)) == tuple(info["config"])
        assert np.array_equal(ref.predict(X), tuned.predict(X))
        for a, b in zip(ref.forest, tuned.forest):
            assert a["leaf_value"] == b["leaf_value"] if "leaf_value" in a else a["bin"] == b["bin"]
        again = WarpGBM(**kw).fit(X, y)
        assert again.hist_config_["cached"] and tuple(again.hist_config_["config"]) == tuple(info["config"])
        w = np.random.default_rng(1).uniform(0.5, 2, len(y)).astype(np.float32)
        wt = WarpGBM(**kw).fit(X, y, sample_weight=w)  # two planes: a shape of its own
        assert not wt.hist_config_["cached"]
        Xi = np.random.default_rng(2).integers(0, 5, (20000, 64)).astype(np.int8)
        m5 = WarpGBM(**dict(kw, num_bins=5)).fit(Xi, y)
        m127 = WarpGBM(**dict(kw, num_bins=127)).fit(X, y)
        assert m5.hist_config_ is not None and m127.hist_config_ is not None
    finally:
        for nb, one in ((32, 1), (32, 0), (5, 1), (127, 1)):
            _lib.call("wgbm_set_hist_config_shape", 64, nb, 64, one, 0, 0, 0, 0, 0)
