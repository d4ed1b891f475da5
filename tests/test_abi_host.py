"""CPU checks of the C ABI library and the host-side logic (no GPU calls).

* every function declared in include/warpgbm_amd.h is exported by the built
  library and bound in warpgbm_amd._lib;
* argument validation in the library returns an error code and message
  before touching the device;
* host logic of the estimator: hyper-parameter validation, monotone
  constraint normalisation, flattening, and the level-order node tables ->
  nested tree dicts conversion.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "warpgbm_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wgbm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from warpgbm_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n
    assert lib.wgbm_build_arch() == b"gfx950"


def test_library_rejects_bad_arguments_without_a_device():
    from warpgbm_amd import _lib
    lib = _lib.load()
    rc = lib.wgbm_compute_split(None, None, 1, 4, 1, 0.0, 1.0, 0.0, None, None, None)
    assert rc != 0 and b"shape" in lib.wgbm_last_error()
    rc = lib.wgbm_set_hist_config(0, 3, 0, 0, 0)
    assert rc != 0
    rc = lib.wgbm_set_hist_config(0, 0, 5, 0, 0)
    assert rc != 0 and b"features_per_lane" in lib.wgbm_last_error()
    rc = lib.wgbm_set_hist_config(0, 0, 0, 0, 300)
    assert rc != 0 and b"slots" in lib.wgbm_last_error()
    with pytest.raises(RuntimeError, match="num_bins"):
        _lib.call("wgbm_compute_histogram3", None, 10, 4, None, None, None, 5, None, 4, None, None, None, 1, 500,
                  None, 0, None)


def test_workspace_queries():
    from warpgbm_amd import _lib
    lib = _lib.load()
    assert lib.wgbm_histogram3_workspace_bytes(1000, 50, 32, 1, 50) > 0
    assert (lib.wgbm_histogram3_workspace_bytes(1000, 50, 32, 7, 50)
            > lib.wgbm_histogram3_workspace_bytes(1000, 50, 32, 1, 50))
    # the bin matrix's column count decides the tiling, hence the workspace:
    # a feature subset of a wide matrix may need more than n_feat columns would
    assert lib.wgbm_histogram3_workspace_bytes(1000, 150, 8, 2, 300) > 0
    c = _lib.FitConfig(n_rows=1_000_000, n_cols=1000, n_feat=1000, n_bins=32, n_eras=1, n_classes=1, max_depth=5,
                       objective=0, learning_rate=0.01, min_split_gain=0.0, min_child_weight=20.0, lambda_=1e-6)
    ws = lib.wgbm_fit_workspace_bytes(ctypes.byref(c))
    assert 40e6 < ws < 2e9  # perm/gh buffers, histograms, partials
    c.n_bins = 1
    assert lib.wgbm_fit_workspace_bytes(ctypes.byref(c)) == 0
    out = (ctypes.c_int32 * 5)()
    lib.wgbm_get_hist_config(1000, 32, 1000, 0, out)
    assert out[0] > 0 and out[1] in (1, 2, 4, 8, 16) and out[2] in (1, 2, 4, 8, 16) and out[3] in (1, 2, 4, 8, 16, 32, 64)
    assert out[4] == 512


def test_hyperparameter_validation_messages():
    from warpgbm_amd import WarpGBM
    with pytest.raises(ValueError, match="num_bins must be between 2 and 127"):
        WarpGBM(num_bins=200)
    with pytest.raises(TypeError, match="max_depth must be an integer"):
        WarpGBM(max_depth=2.5)
    with pytest.raises(ValueError, match="objective must be"):
        WarpGBM(objective="poisson")
    with pytest.raises(ValueError, match="engine"):
        WarpGBM(engine="cpu")
    with pytest.raises(TypeError, match="monotonic_constraints must be None"):
        WarpGBM(monotonic_constraints="up")


def test_monotone_normalisation():
    from warpgbm_amd import WarpGBM
    m = WarpGBM(monotonic_constraints={2: 1, 0: -1})
    assert m._normalize_monotonic_constraints(4).tolist() == [-1, 0, 1, 0]
    m = WarpGBM(monotonic_constraints=[1, 0])
    with pytest.raises(ValueError, match="length must equal n_features"):
        m._normalize_monotonic_constraints(3)
    m = WarpGBM(monotonic_constraints={5: 1})
    with pytest.raises(ValueError, match="out of range"):
        m._normalize_monotonic_constraints(3)


def test_era_inverse_equals_numpy_unique():
    """The counting era map gives np.unique's (sorted ids, inverse) pair for
    constant, narrow, negative, gapped, wide (sorted fallback), empty and
    non-integer era arrays."""
    from warpgbm_amd.core import _era_inverse
    rng = np.random.default_rng(0)
    cases = [np.ones(1000, np.int32), np.zeros(7, np.int64), rng.integers(-50, 50, 5000),
             rng.integers(0, 1 << 30, 3000), rng.choice([3, 17, 99, -4], 10000), np.array([5]),
             np.array([], np.int32), rng.integers(0, 200, 100_000).astype(np.int64),
             np.array([2**31 - 1, -2**31, 0], np.int64), rng.normal(size=100) * 10]
    for c in cases:
        u, inv = _era_inverse(c)
        u_ref, inv_ref = np.unique(np.asarray(c).astype(np.int32), return_inverse=True)
        assert u.dtype == u_ref.dtype and np.array_equal(u, u_ref)
        assert np.array_equal(inv.reshape(-1), inv_ref.reshape(-1))


def test_fit_argument_errors_before_any_device_work():
    """tests/test_sample_weight.py:121-139 and
    tests/test_monotonic_constraints.py:194-240: bad weights and bad
    constraints are refused by fit with the reference's messages, before
    the data reaches a device (so this runs without a GPU)."""
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(3)
    X = rng.standard_normal((120, 3)).astype(np.float32)
    y = rng.standard_normal(120).astype(np.float32)
    m = WarpGBM(n_estimators=5, max_depth=2)
    w = np.ones(120, np.float32)
    w[0] = 0.0
    with pytest.raises(ValueError, match="strictly positive"):
        m.fit(X, y, sample_weight=w)
    with pytest.raises(ValueError, match="same length"):
        m.fit(X, y, sample_weight=np.ones(119, np.float32))
    with pytest.raises(ValueError, match="sample_weight_eval can only be used"):
        m.fit(X, y, sample_weight_eval=np.ones(120, np.float32))
    cases = [([1, 0], ValueError, "length must equal n_features"),
             ([1, 2, 0], ValueError, "values must be one of"),
             ({5: 1}, ValueError, "out of range"),
             ({"0": 1}, TypeError, "integer feature indices")]
    for mono, err, msg in cases:
        with pytest.raises(err, match=msg):
            WarpGBM(n_estimators=5, max_depth=2, monotonic_constraints=mono).fit(X, y)


def test_tables_to_dicts_and_flatten_roundtrip():
    """Level-order node tables (as the batched trainer writes them) become
    the reference's nested dicts, whose flattening traverses identically."""
    import torch
    from oracle import gbm_oracle as O
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(0)
    N, F, B, D = 500, 6, 8, 3
    codes = rng.integers(0, B, size=(N, F)).astype(np.int8)
    rec = O.grow_tree(codes, rng.standard_normal(N) + (codes[:, 2] > 3), np.ones(N), np.zeros(N, int), 1, B, D,
                      20, 0.0, 1e-6)
    NH = 2 ** (D + 1) - 1
    st = np.zeros((1, 1, NH), np.int32)
    ft = np.full((1, 1, NH), -1, np.int32)
    bt = np.full((1, 1, NH), -1, np.int32)
    vt = np.zeros((1, 1, NH), np.float32)
    ct = np.zeros((1, 1, NH), np.int32)
    eg = np.zeros((1, 1, NH, 1), np.float32)
    for h, node in rec.nodes.items():
        if node[0] == "split":
            st[0, 0, h], ft[0, 0, h], bt[0, 0, h], eg[0, 0, h, 0] = 1, node[1], node[2], node[3][0]
        else:
            st[0, 0, h], vt[0, 0, h], ct[0, 0, h] = 2, node[1], node[2]
    m = WarpGBM(max_depth=D, n_estimators=1)
    m.objective = "regression"
    m._feature_ids = torch.arange(F, dtype=torch.int32)
    m.per_era_feature_importance_ = np.zeros((1, F), np.float32)
    m.forest = [{}]
    m._trees_trained = 0
    m._append_rounds({"status": st, "feat": ft, "bin": bt, "value": vt, "count": ct, "era_gain": eg}, 1, 1, None)
    t = m.forest[0]
    flat = m.flatten_tree(t, 2 ** (D + 1)).numpy()
    ref = O.predict_tree(rec.tree, codes).astype(np.float32)
    got = O.predict_forest_flat(codes, [flat], 1.0)
    np.testing.assert_allclose(got, ref, rtol=1e-6)
    assert m._feature_of(t["feature"]) == rec.tree["feature"]
    assert m.per_era_feature_importance_[0, rec.tree["feature"]] > 0


def test_tables_to_dicts_many_trees_and_importance_order():
    """n rounds x K classes of random heap tables: every tree equals a
    recursive reading of its heap (split: feature/bin/children, leaf:
    value/rows, absent slots unused), and per-era importances are the split
    gains added in float32 round by round, class by class, in depth-first
    order (the order the per-node path adds them)."""
    import torch
    from warpgbm_amd import WarpGBM
    rng = np.random.default_rng(4)
    n, K, D, F, E = 5, 3, 6, 9, 2
    NH = 2 ** (D + 1) - 1
    st = np.zeros((n, K, NH), np.int32)
    ft = rng.integers(0, F, (n, K, NH)).astype(np.int32)
    bt = rng.integers(0, 31, (n, K, NH)).astype(np.int32)
    vt = rng.standard_normal((n, K, NH)).astype(np.float32)
    ct = rng.integers(1, 999, (n, K, NH)).astype(np.int32)
    eg = (rng.standard_normal((n, K, NH, E)) * 1e3).astype(np.float32)

    def grow(r, c, h, d):
        if d < D and rng.random() < 0.75:
            st[r, c, h] = 1
            grow(r, c, 2 * h + 1, d + 1)
            grow(r, c, 2 * h + 2, d + 1)
        else:
            st[r, c, h] = 2

    for r in range(n):
        for c in range(K):
            grow(r, c, 0, 0)
    imp = np.zeros((E, F), np.float32)

    def read(r, c, h):
        if st[r, c, h] == 1:
            imp[:, ft[r, c, h]] += eg[r, c, h]
            return (int(ft[r, c, h]), int(bt[r, c, h]), read(r, c, 2 * h + 1), read(r, c, 2 * h + 2))
        return (float(vt[r, c, h]), int(ct[r, c, h]))

    want = [[read(r, c, 0) for c in range(K)] for r in range(n)]
    m = WarpGBM(max_depth=D, n_estimators=n)
    m.objective = "multiclass"
    m._feature_ids = torch.arange(F, dtype=torch.int32)
    m.per_era_feature_importance_ = np.zeros((E, F), np.float32)
    m.forest = []
    m._trees_trained = 0
    m._append_rounds({"status": st, "feat": ft, "bin": bt, "value": vt, "count": ct, "era_gain": eg}, n, K, None)

    def as_tuple(t):
        if "leaf_value" in t:
            return (t["leaf_value"], t["samples"])
        return (m._feature_of(t["feature"]), t["bin"], as_tuple(t["left"]), as_tuple(t["right"]))

    assert [[as_tuple(t) for t in trees] for trees in m.forest] == want
    np.testing.assert_array_equal(m.per_era_feature_importance_, imp)


def test_engine_routing_by_depth_and_classes():
    """The batched engine takes every tree its level tables hold (depth <= 20
    and classes x 2^depth < 2^24); deeper trees and engine="plain" run per
    node."""
    from warpgbm_amd import WarpGBM
    m = WarpGBM(max_depth=20)
    m.objective, m.num_classes = "regression", None
    assert not m._use_plain()
    m.max_depth = 21
    assert m._use_plain()
    m = WarpGBM(max_depth=20, objective="multiclass")
    m.num_classes = 16  # 16 << 20 == 2^24
    assert m._use_plain()
    m.num_classes = 15
    assert not m._use_plain()
    assert WarpGBM(max_depth=3, engine="plain")._use_plain()


def test_engine_routing_by_class_count_rows_and_memory(monkeypatch):
    """More than 256 classes, classes x rows of 2^31 or more, or a batched
    workspace above 80 % of the device's free memory route the fit to the
    per-node engine (the batched engine's limits; the per-node engine has
    none of them); node tables are written in calls of bounded size."""
    import torch
    from warpgbm_amd import WarpGBM, _lib
    m = WarpGBM(objective="multiclass", max_depth=4, num_bins=32)
    m.num_classes = 300
    assert m._use_plain()
    m.num_classes = 5
    assert not m._use_plain()
    m.num_samples, m.num_features, m.num_eras = 500_000, 2000, 500
    m.max_depth = 10
    m.num_bins = 5
    assert not m._use_plain()  # no device to ask: the batched engine stays
    need = _lib.query("wgbm_fit_workspace_bytes", ctypes.byref(m._fit_config(m._n_sampled(), 5)))
    assert need > 10 << 30
    monkeypatch.setattr(torch.cuda, "is_available", lambda: True)
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (int(need / 0.8) + 1, 288 << 30))
    m.device = "cuda:0"
    assert not m._use_plain()
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (int(need / 0.8) - 1, 288 << 30))
    assert m._use_plain()
    m.num_samples = (1 << 31) // 5 + 1
    assert m._use_plain()
    # node tables: 1000 rounds of depth-12, 500-era trees go in several calls
    m.max_depth, m.num_eras = 12, 500
    per_round = 5 * (2 ** 13 - 1) * (20 + 4 * 500)
    assert m._rounds_per_call(5) == (1 << 30) // per_round
    m1 = WarpGBM(max_depth=3)
    m1.num_eras = 1
    assert m1._rounds_per_call(1) > 1000


def test_server_refuses_requests_after_close():
    """submit() after close() (or after the worker thread died) raises
    instead of queueing a request nothing would serve."""
    import collections
    import threading
    from types import SimpleNamespace
    from warpgbm_amd.server import PredictServer
    srv = PredictServer.__new__(PredictServer)  # no device: the worker and streams are stand-ins
    srv.forest = SimpleNamespace(n_feat=4, n_classes=1)
    srv.max_batch_rows, srv.budget, srv.bytes_per_row = 100, 1 << 20, 16
    srv._cv = threading.Condition()
    srv._queue = collections.deque()
    srv._stop = False
    srv._thread = threading.Thread(target=lambda: None)
    srv._thread.start()
    srv._thread.join()  # a worker that has ended
    with pytest.raises(RuntimeError, match="closed"):
        srv.submit(np.zeros((3, 4), np.int8))
    srv._stop = True
    with pytest.raises(RuntimeError, match="closed"):
        srv.submit(np.zeros((3, 4), np.int8))
    assert not srv._queue


def test_reference_import_names_through_the_package_alias():
    """A caller written against the reference (``from warpgbm import
    WarpGBM``, ``from warpgbm.cuda import node_kernel``) runs unchanged once
    the package is aliased as ``warpgbm``; run in a child interpreter so the
    alias does not leak into other tests."""
    import subprocess
    import sys
    code = (
        "import sys, warpgbm_amd; sys.modules['warpgbm'] = warpgbm_amd\n"
        "from warpgbm import WarpGBM\n"
        "from warpgbm.cuda import node_kernel\n"
        "from warpgbm.metrics import weighted_mse_torch, weighted_multi_mse_torch\n"
        "assert WarpGBM is warpgbm_amd.WarpGBM\n"
        "for name in ('compute_histogram3', 'compute_split', 'custom_cuda_binner', 'predict_forest'):\n"
        "    assert callable(getattr(node_kernel, name)), name\n"
        "m = WarpGBM(objective='multi_regression', n_estimators=3, device='cpu')\n"
        "assert m.objective == 'multi_regression'\n"
        "print('ok')\n")
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]
