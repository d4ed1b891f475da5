"""WarpGBM estimator on the MI355X engine.

Drop-in for ``warpgbm.core.WarpGBM`` (warpgbm/core.py:57-2328): the same
constructor, ``fit`` / ``predict`` / ``predict_proba`` behaviour, attributes
and saved-model dict.  All device work goes through the gfx950 HIP library
(``libwarpgbm_amd.so``):

* ``engine="batched"`` (default) trains with ``wgbm_fit_rounds``: every
  frontier node x era x class of a tree level in one histogram launch, the
  DES scan fused behind it, partition and leaf updates on device, no host
  synchronisation inside a training call.
* ``engine="plain"`` keeps the reference's per-node structure
  (grow_tree recursion, core.py:1364-1470): one histogram / split / select /
  partition / leaf launch per node with the host syncs that structure
  implies.  It is the baseline the batched engine is timed against; both
  engines produce identical trees.

The extension never falls back to CPU code: without the built library every
fit/predict raises ImportError.
"""
from __future__ import annotations

import ctypes
import pickle
import time

import numpy as np
import torch
from sklearn.base import BaseEstimator, RegressorMixin
from sklearn.preprocessing import LabelEncoder

from . import _lib
from .cuda import node_kernel as _nk
from .metrics import (
    softmax,
    weighted_accuracy_torch,
    weighted_corr_loss_torch,
    weighted_log_loss_torch,
    weighted_mse_torch,
    weighted_multi_mse_torch,
    weighted_rmsle_torch,
)

_OBJECTIVES = ["regression", "multi_regression", "multiclass", "binary"]
_ENGINES = ("batched", "plain")


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class _Workspace:
    """Grow-only scratch buffers on one device (C-ABI workspaces)."""

    def __init__(self):
        self._bufs = {}

    def get(self, key, nbytes, device):
        nbytes = max(int(nbytes), 256)
        b = self._bufs.get(key)
        if b is None or b.numel() < nbytes or b.device != device:
            b = torch.empty(nbytes, dtype=torch.uint8, device=device)
            self._bufs[key] = b
        return b


# The batched engine's fit workspace (the column-major code copy, level
# histograms, tables) is scratch for the duration of one wgbm_fit_rounds call:
# the call returns when its rounds are done.  One buffer per (device, stream)
# therefore serves every model in turn, so a model fitted while an earlier one
# is still alive reuses the buffer instead of allocating (and first touching)
# gigabytes of its own, and no model keeps it after its fit.  Fits on one
# device and stream run one after another.
_FIT_WORKSPACES: dict = {}


def _fit_ws_key(device):
    d = torch.device(device)
    return (d.index if d.index is not None else torch.cuda.current_device(), _stream(d))


def _fit_workspace(device, nbytes: int) -> torch.Tensor:
    key = _fit_ws_key(device)
    b = _FIT_WORKSPACES.get(key)
    if b is None or b.numel() < nbytes:
        _FIT_WORKSPACES.pop(key, None)  # the old buffer goes back to the allocator first
        b = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
        _FIT_WORKSPACES[key] = b
    return b


_STAGE_BYTES = 64 << 20  # pinned staging chunk for large host-to-device copies


def _to_device(X_np, device, dtype=torch.float32):
    """Copy of a host array on the device (converted to dtype on the host
    first when given).  Arrays of four or more staging chunks go through two
    pinned buffers in chunks of _STAGE_BYTES, so the host-side copy of one
    chunk overlaps the DMA of the previous one (a pageable copy stages
    through the driver's buffers one at a time)."""
    device = torch.device(device)
    src = torch.from_numpy(np.ascontiguousarray(X_np))
    if dtype is not None and src.dtype != dtype:
        src = src.to(dtype)
    es = src.element_size()
    if device.type != "cuda" or src.numel() * es < 4 * _STAGE_BYTES:
        return src.to(device)
    out = torch.empty(src.shape, dtype=src.dtype, device=device)
    flat_src, flat_out = src.reshape(-1), out.reshape(-1)
    step = max(_STAGE_BYTES // es, 1)
    bufs = [torch.empty(step, dtype=src.dtype, pin_memory=True) for _ in range(2)]
    done = [None, None]
    stream = torch.cuda.current_stream(device)
    for i, off in enumerate(range(0, flat_src.numel(), step)):
        k = i & 1
        if done[k] is not None:
            done[k].synchronize()  # the buffer's previous DMA has finished
        m = min(step, flat_src.numel() - off)
        bufs[k][:m].copy_(flat_src[off:off + m])
        flat_out[off:off + m].copy_(bufs[k][:m], non_blocking=True)
        done[k] = torch.cuda.Event()
        done[k].record(stream)
    for e in done:
        if e is not None:
            e.synchronize()  # the staging buffers are released below
    return out


def _era_inverse(era_id):
    """(sorted unique era ids, index of each row's id among them), as
    np.unique(era_id.astype(int32), return_inverse=True) gives them
    (warpgbm/core.py:1284 takes torch.unique's, the same pair).  Ids that
    are all equal, or span fewer than 2^22 values, are mapped by counting
    in O(n) instead of sorting (a million rows: ~1 ms against 4 ms sorted,
    and ~90 ms for the first sort of a process)."""
    e = np.asarray(era_id).astype(np.int32).reshape(-1)
    if e.size == 0:
        return np.unique(e, return_inverse=True)
    lo, hi = int(e.min()), int(e.max())
    if lo == hi:
        return np.array([lo], np.int32), np.zeros(e.size, np.int64)
    if hi - lo < (1 << 22):
        off = (e - lo).astype(np.int64) if lo != 0 else e.astype(np.int64)
        present = np.bincount(off, minlength=hi - lo + 1) > 0
        rank = np.cumsum(present) - 1
        return (np.flatnonzero(present) + lo).astype(np.int32), rank[off]
    return np.unique(e, return_inverse=True)


_PREORDER = {}


def _heap_preorder(n_slots):
    """Heap slots 0..n_slots-1 in depth-first preorder (left child first)."""
    p = _PREORDER.get(n_slots)
    if p is None:
        out, stack = [], [0]
        while stack:
            h = stack.pop()
            out.append(h)
            if 2 * h + 2 < n_slots:
                stack.append(2 * h + 2)
                stack.append(2 * h + 1)
        p = _PREORDER[n_slots] = np.asarray(out, dtype=np.int64)
    return p


class WarpGBM(BaseEstimator, RegressorMixin):
    def __init__(
        self,
        objective="regression",
        num_bins=10,
        max_depth=3,
        learning_rate=0.1,
        n_estimators=100,
        min_child_weight=20,
        min_split_gain=0.0,
        threads_per_block=64,
        rows_per_thread=4,
        L2_reg=1e-6,
        device="cuda",
        colsample_bytree=1.0,
        random_state=None,
        warm_start=False,
        monotonic_constraints=None,
        engine="batched",
        autotune=True,
    ):
        self._validate_hyperparams(
            objective=objective, num_bins=num_bins, max_depth=max_depth, learning_rate=learning_rate,
            n_estimators=n_estimators, min_child_weight=min_child_weight, min_split_gain=min_split_gain,
            threads_per_block=threads_per_block, rows_per_thread=rows_per_thread, L2_reg=L2_reg,
            colsample_bytree=colsample_bytree, monotonic_constraints=monotonic_constraints,
        )
        if engine not in _ENGINES:
            raise ValueError(f"engine must be one of {_ENGINES}, got {engine!r}.")
        self.objective = objective
        self.num_bins = num_bins
        self.max_depth = max_depth
        self.learning_rate = learning_rate
        self.n_estimators = n_estimators
        self.min_child_weight = min_child_weight
        self.min_split_gain = min_split_gain
        self.threads_per_block = threads_per_block
        self.rows_per_thread = rows_per_thread
        self.L2_reg = L2_reg
        self.device = device
        self.colsample_bytree = colsample_bytree
        self.random_state = random_state
        self.warm_start = warm_start
        self.monotonic_constraints = monotonic_constraints
        self.engine = engine
        # The batched engine times the histogram tilings on the first fit of
        # each shape (a few rounds on the fit's own data, cached per shape) and
        # trains with the fastest; it takes over from threads_per_block /
        # rows_per_thread, which this engine accepts and ignores.
        self.autotune = autotune
        self.hist_config_ = None
        self.forest = [{} for _ in range(self.n_estimators)]
        self.bin_edges = None
        self.base_prediction = None
        self.unique_eras = None
        self.num_classes = None
        self.num_outputs = None
        self.classes_ = None
        self.num_features = None
        self.num_samples = None
        self.label_encoder = None
        self.feature_importance_ = None
        self.per_era_feature_importance_ = None
        self.monotonic_constraints_vector_ = None
        self.monotonic_required_split_direction_ = None
        self.monotonic_required_split_direction_gpu_ = None
        self._monotonic_constraints_fitted_ = False
        self._is_fitted = False
        self._trees_trained = 0
        self.training_loss = []
        self.eval_loss = []
        self._ws = _Workspace()
        self._feature_ids = None
        self._invalidate_prediction_cache()

    # ------------------------------------------------------------ validation
    def _validate_hyperparams(self, **kw):
        """Hyper-parameter checks of warpgbm/core.py:147-219."""
        if kw["objective"] not in _OBJECTIVES:
            raise ValueError("objective must be 'regression', 'multi_regression', 'binary', "
                             f"or 'multiclass', got {kw['objective']}.")
        for p in ["num_bins", "max_depth", "n_estimators", "min_child_weight", "threads_per_block",
                  "rows_per_thread"]:
            if not isinstance(kw[p], int):
                raise TypeError(f"{p} must be an integer, got {type(kw[p])}.")
        for p in ["learning_rate", "min_split_gain", "L2_reg", "colsample_bytree"]:
            if not isinstance(kw[p], (float, int)):
                raise TypeError(f"{p} must be a float, got {type(kw[p])}.")
        if not (2 <= kw["num_bins"] <= 127):
            raise ValueError("num_bins must be between 2 and 127 inclusive.")
        if kw["max_depth"] < 1:
            raise ValueError("max_depth must be at least 1.")
        if not (0.0 < kw["learning_rate"] <= 1.0):
            raise ValueError("learning_rate must be in (0.0, 1.0].")
        if kw["n_estimators"] <= 0:
            raise ValueError("n_estimators must be positive.")
        if kw["min_child_weight"] < 1:
            raise ValueError("min_child_weight must be a positive integer.")
        if kw["min_split_gain"] < 0:
            raise ValueError("min_split_gain must be non-negative.")
        if kw["threads_per_block"] <= 0 or kw["threads_per_block"] % 32 != 0:
            raise ValueError("threads_per_block should be a positive multiple of 32 (warp size).")
        if not (1 <= kw["rows_per_thread"] <= 16):
            raise ValueError("rows_per_thread must be positive between 1 and 16 inclusive.")
        if kw["L2_reg"] < 0:
            raise ValueError("L2_reg must be non-negative.")
        if kw["colsample_bytree"] <= 0 or kw["colsample_bytree"] > 1:
            raise ValueError(f"Invalid colsample_bytree: {kw['colsample_bytree']}. "
                             "Must be a float value > 0 and <= 1.")
        mc = kw.get("monotonic_constraints")
        if mc is not None and not isinstance(mc, (list, tuple, np.ndarray, dict)):
            raise TypeError("monotonic_constraints must be None, a vector-like object, or a dict "
                            "{feature_idx: -1|0|1}.")

    def _normalize_monotonic_constraints(self, n_features):
        """Dense int8 constraint vector, warpgbm/core.py:221-256."""
        c = self.monotonic_constraints
        if c is None:
            return np.zeros(n_features, dtype=np.int8)
        if isinstance(c, dict):
            dense = np.zeros(n_features, dtype=np.int8)
            for key, value in c.items():
                if isinstance(key, bool) or not isinstance(key, (int, np.integer)):
                    raise TypeError("monotonic_constraints dict keys must be integer feature indices.")
                if key < 0 or key >= n_features:
                    raise ValueError(f"monotonic_constraints key {key} out of range for {n_features} features.")
                if value not in (-1, 0, 1):
                    raise ValueError("monotonic_constraints values must be one of {-1, 0, 1}.")
                dense[int(key)] = int(value)
            return dense
        dense = np.asarray(c, dtype=np.int8)
        if dense.ndim != 1:
            raise ValueError(f"monotonic_constraints vector must be 1-dimensional, got shape {dense.shape}.")
        if dense.shape[0] != n_features:
            raise ValueError("monotonic_constraints vector length must equal n_features "
                             f"({n_features}), got {dense.shape[0]}.")
        if not np.isin(dense, [-1, 0, 1]).all():
            raise ValueError("monotonic_constraints values must be one of {-1, 0, 1}.")
        return dense.astype(np.int8, copy=False)

    def _prepare_monotonic_constraints(self, n_features):
        """Required split direction = -constraint (warpgbm/core.py:258-276)."""
        dense = self._normalize_monotonic_constraints(n_features)
        if self.warm_start and self._is_fitted and self._monotonic_constraints_fitted_:
            if not np.array_equal(dense, self.monotonic_constraints_vector_):
                raise ValueError("monotonic_constraints must match the fitted model when warm_start=True.")
        self.monotonic_constraints_vector_ = dense
        self.monotonic_required_split_direction_ = -dense
        self.monotonic_required_split_direction_gpu_ = (
            torch.from_numpy(np.ascontiguousarray(self.monotonic_required_split_direction_)).to(torch.int8)
            .to(self.device))
        self._monotonic_constraints_fitted_ = True

    def validate_fit_params(self, X, y, era_id, X_eval, y_eval, eval_every_n_trees, early_stopping_rounds,
                            eval_metric, sample_weight, sample_weight_eval):
        """Argument checks of warpgbm/core.py:369-558; returns early_stopping_rounds."""
        multi = self.objective == "multi_regression"
        if not isinstance(X, np.ndarray) or not isinstance(y, np.ndarray):
            raise TypeError("X and y must be numpy arrays.")
        if X.ndim != 2:
            raise ValueError(f"X must be 2-dimensional, got shape {X.shape}")
        if multi:
            if y.ndim != 2:
                raise ValueError(f"y must be 2-dimensional for objective='multi_regression', got shape {y.shape}")
            if y.shape[1] < 1:
                raise ValueError("y must contain at least one output column.")
        elif y.ndim != 1:
            raise ValueError(f"y must be 1-dimensional, got shape {y.shape}")
        if X.shape[0] != y.shape[0]:
            raise ValueError(f"X and y must have the same number of rows. Got {X.shape[0]} and {y.shape[0]}.")
        if sample_weight is not None:
            self._check_weight(sample_weight, y, "sample_weight", multi)
        if sample_weight_eval is not None and (X_eval is None or y_eval is None):
            raise ValueError("sample_weight_eval can only be used when X_eval and y_eval are provided.")
        if era_id is not None:
            if not isinstance(era_id, np.ndarray):
                raise TypeError("era_id must be a numpy array.")
            if era_id.ndim != 1:
                raise ValueError(f"era_id must be 1-dimensional, got shape {era_id.shape}")
            if len(era_id) != X.shape[0]:
                raise ValueError(f"era_id must have same length as y. Got {len(era_id)} and {X.shape[0]}.")
        if any(a is not None for a in (X_eval, y_eval, eval_every_n_trees)):
            if X_eval is None or y_eval is None or eval_every_n_trees is None:
                raise ValueError("If using eval set, X_eval, y_eval, and eval_every_n_trees must all be defined.")
            if not isinstance(X_eval, np.ndarray) or not isinstance(y_eval, np.ndarray):
                raise TypeError("X_eval and y_eval must be numpy arrays.")
            if X_eval.ndim != 2:
                raise ValueError(f"X_eval must be 2-dimensional, got shape {X_eval.shape}")
            if multi:
                if y_eval.ndim != 2:
                    raise ValueError("y_eval must be 2-dimensional for objective='multi_regression', "
                                     f"got shape {y_eval.shape}")
                if y_eval.shape[1] != y.shape[1]:
                    raise ValueError("y_eval must have the same number of outputs as y. "
                                     f"Got {y_eval.shape[1]} and {y.shape[1]}.")
            elif y_eval.ndim != 1:
                raise ValueError(f"y_eval must be 1-dimensional, got shape {y_eval.shape}")
            if X_eval.shape[0] != y_eval.shape[0]:
                raise ValueError("X_eval and y_eval must have same number of rows. "
                                 f"Got {X_eval.shape[0]} and {y_eval.shape[0]}.")
            if sample_weight_eval is not None:
                self._check_weight(sample_weight_eval, y_eval, "sample_weight_eval", multi)
            if not isinstance(eval_every_n_trees, int) or eval_every_n_trees <= 0:
                raise ValueError(f"eval_every_n_trees must be a positive integer, got {eval_every_n_trees}.")
            if early_stopping_rounds is not None:
                if not isinstance(early_stopping_rounds, int) or early_stopping_rounds <= 0:
                    raise ValueError(f"early_stopping_rounds must be a positive integer, got {early_stopping_rounds}.")
            else:
                early_stopping_rounds = self.n_estimators + 1
            valid = ["mse"] if multi else ["mse", "corr", "rmsle", "logloss", "accuracy"]
            if eval_metric not in valid:
                raise ValueError(f"Invalid eval_metric: {eval_metric}. Choose from {valid}.")
        return early_stopping_rounds

    def _check_weight(self, w, y, name, multi):
        if not isinstance(w, np.ndarray):
            raise TypeError(f"{name} must be a numpy array.")
        if multi:
            n, k = y.shape
            if w.ndim == 1:
                if w.shape[0] != n:
                    raise ValueError(f"{name} must have same length as y rows. Got {w.shape[0]} and {n}.")
            elif w.ndim == 2:
                if w.shape != (n, k):
                    raise ValueError(f"{name} must have shape {(n, k)} for multi-output targets, got {w.shape}.")
            else:
                raise ValueError(f"{name} must be 1-dimensional or 2-dimensional, got shape {w.shape}.")
        else:
            if w.ndim != 1:
                raise ValueError(f"{name} must be 1-dimensional, got shape {w.shape}")
            if len(w) != len(y):
                raise ValueError(f"{name} must have same length as {'y' if name == 'sample_weight' else 'y_eval'}. "
                                 f"Got {len(w)} and {len(y)}.")
        if not np.all(np.isfinite(w)):
            raise ValueError(f"{name} must contain only finite values.")
        if multi:
            if not np.all(w >= 0):
                raise ValueError(f"{name} must be non-negative.")
            if not np.any(w > 0):
                raise ValueError(f"{name} must contain at least one positive value.")
        elif not np.all(w > 0):
            raise ValueError(f"{name} must be strictly positive.")

    def _normalize_sample_weight(self, w):
        """Mean-normalised float32 weights, warpgbm/core.py:589-594."""
        w = w.astype(np.float32, copy=False)
        m = float(w.mean())
        if m <= 0:
            raise ValueError("sample weights must have positive mean.")
        return w / m

    # ------------------------------------------------------------ caches
    def _invalidate_prediction_cache(self):
        self._heap_cache = None
        self._cached_regression_tree_tensor = None
        self._cached_multi_regression_tree_tensors = None
        self._cached_multiclass_tree_tensors = None
        self._cached_tree_device = None
        self._cached_tree_max_nodes = None
        self._cached_tree_objective = None
        self._cached_tree_rounds = None
        self._cached_tree_num_outputs = None

    def _prediction_cache_matches(self, key, max_nodes, num_outputs=None):
        return (self._cached_tree_device == str(self.device) and self._cached_tree_max_nodes == max_nodes
                and self._cached_tree_objective == key and self._cached_tree_rounds == self._trees_trained
                and self._cached_tree_num_outputs == num_outputs)

    def _set_prediction_cache_meta(self, key, max_nodes, num_outputs=None):
        self._cached_tree_device = str(self.device)
        self._cached_tree_max_nodes = max_nodes
        self._cached_tree_objective = key
        self._cached_tree_rounds = self._trees_trained
        self._cached_tree_num_outputs = num_outputs

    def _get_cached_regression_tree_tensor(self):
        """Flattened forest [T, 2^(D+1), 6], cached per tree count (core.py:622-647)."""
        max_nodes = 2 ** (self.max_depth + 1)
        if (self._cached_regression_tree_tensor is not None
                and self._prediction_cache_matches("regression", max_nodes, num_outputs=1)):
            return self._cached_regression_tree_tensor
        trees = [t for t in self.forest[: self._trees_trained] if t]
        if not trees:
            trees = [t for t in self.forest if t]
        if trees:
            tt = torch.stack([self.flatten_tree(t, max_nodes=max_nodes) for t in trees]).to(self.device)
        else:
            tt = torch.empty((0, max_nodes, 6), dtype=torch.float32, device=self.device)
        self._cached_regression_tree_tensor = tt
        self._cached_multi_regression_tree_tensors = None
        self._cached_multiclass_tree_tensors = None
        self._set_prediction_cache_meta("regression", max_nodes, num_outputs=1)
        return tt

    def _get_cached_multiclass_tree_tensors(self):
        """Per-class flattened forests (core.py:649-686)."""
        max_nodes = 2 ** (self.max_depth + 1)
        if (self._cached_multiclass_tree_tensors is not None
                and self._prediction_cache_matches("classification", max_nodes, num_outputs=self.num_classes)):
            return self._cached_multiclass_tree_tensors
        lists = [[] for _ in range(self.num_classes)]
        for round_trees in self.forest[: self._trees_trained]:
            if not round_trees:
                continue
            for k, t in enumerate(round_trees):
                if t:
                    lists[k].append(self.flatten_tree(t, max_nodes=max_nodes))
        out = [torch.stack(l).to(self.device) if l else None for l in lists]
        self._cached_multiclass_tree_tensors = out
        self._cached_regression_tree_tensor = None
        self._cached_multi_regression_tree_tensors = None
        self._set_prediction_cache_meta("classification", max_nodes, num_outputs=self.num_classes)
        return out

    # ------------------------------------------------------------ fit
    def fit(self, X, y, era_id=None, sample_weight=None, X_eval=None, y_eval=None, sample_weight_eval=None,
            eval_every_n_trees=None, early_stopping_rounds=None, eval_metric="mse"):
        """Train the forest (warpgbm/core.py:733-881)."""
        early_stopping_rounds = self.validate_fit_params(
            X, y, era_id, X_eval, y_eval, eval_every_n_trees, early_stopping_rounds, eval_metric,
            sample_weight, sample_weight_eval)
        multi = self.objective == "multi_regression"
        # multi-output weights default to ones of y's shape (core.py:759-772)
        sample_weight = ((np.ones_like(y, dtype=np.float32) if multi else np.ones(X.shape[0], dtype=np.float32))
                         if sample_weight is None else self._normalize_sample_weight(sample_weight))
        if X_eval is not None and y_eval is not None:
            sample_weight_eval = ((np.ones_like(y_eval, dtype=np.float32) if multi
                                   else np.ones(X_eval.shape[0], dtype=np.float32))
                                  if sample_weight_eval is None else self._normalize_sample_weight(sample_weight_eval))
        if self.random_state is not None:
            torch.manual_seed(self.random_state)
            np.random.seed(self.random_state)
            if torch.cuda.is_available():
                torch.cuda.manual_seed_all(self.random_state)
        if not self.warm_start or not self._is_fitted:
            self._is_fitted = False
            self._trees_trained = 0
            self.forest = ([{} for _ in range(self.n_estimators)]
                           if self.objective in ("regression", "multi_regression") else [])
            self.training_loss = []
            self.eval_loss = []
            self._round_tables = []
            self._invalidate_prediction_cache()
        else:
            if X.shape[1] != self.num_features:
                raise ValueError(f"X has {X.shape[1]} features, but model was trained with "
                                 f"{self.num_features} features.")
            if self._trees_trained >= self.n_estimators:
                print(f"Model already has {self._trees_trained} trees (n_estimators={self.n_estimators}). "
                      "No additional training needed.")
                return self
        if era_id is None:
            era_id = np.ones(X.shape[0], dtype="int32")
        self._prepare_monotonic_constraints(X.shape[1])
        if self.objective in ("multiclass", "binary"):
            if not self.warm_start or not self._is_fitted:
                self.label_encoder = LabelEncoder()
                y_enc = self.label_encoder.fit_transform(y)
                self.classes_ = self.label_encoder.classes_
                self.num_classes = len(self.classes_)
            else:
                y_enc = self.label_encoder.transform(y)
            if self.objective == "binary" and self.num_classes != 2:
                raise ValueError(f"binary objective requires exactly 2 classes, got {self.num_classes}")
            if self.objective == "multiclass" and self.num_classes < 2:
                raise ValueError(f"multiclass objective requires at least 2 classes, got {self.num_classes}")
            return self._fit_classification(X, y_enc, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                                            eval_every_n_trees, early_stopping_rounds, eval_metric)
        if multi:
            return self._fit_multi_regression(X, y, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                                              eval_every_n_trees, early_stopping_rounds, eval_metric)
        return self._fit_regression(X, y, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                                    eval_every_n_trees, early_stopping_rounds, eval_metric)

    def _setup_common(self, X, y_float, era_id, sample_weight):
        (self.bin_indices, self.era_indices, self.bin_edges, self.unique_eras, self.Y_gpu) = \
            self.preprocess_gpu_data(X, y_float, era_id)
        self.num_samples, self.num_features = X.shape
        self.num_eras = len(self.unique_eras)
        self.sample_weight_gpu = torch.from_numpy(np.ascontiguousarray(sample_weight, np.float32)).to(self.device)
        self._const_weight = None
        self._feature_ids = torch.arange(self.num_features, dtype=torch.int32, device=self.device)
        self.feature_indices = self._feature_ids
        # Root row order: rows grouped by era, stable (identity for one era).
        era_np = self._era_np
        if self.num_eras > 1:
            order = np.argsort(era_np, kind="stable").astype(np.int32)
            self._root_order = torch.from_numpy(order).to(self.device)
        else:
            self._root_order = None
        self._era_count = torch.from_numpy(
            np.bincount(era_np, minlength=self.num_eras).astype(np.int32)).to(self.device)
        self.root_node_indices = (self._root_order if self._root_order is not None else
                                  torch.arange(self.num_samples, dtype=torch.int32, device=self.device))

    def fit_binned(self, bin_indices, y, era_id=None, sample_weight=None, bin_edges=None):
        """Regression fit from int8 bin codes already resident on the device.

        ``bin_indices``: int8 [N, F] device tensor; ``y``: float32 [N] device
        tensor.  No host copies and no binning: this is the training step
        of ``fit`` on data that already lives in HBM (benchmarks, pipelines
        that bin once and train many models).  ``bin_edges`` (float32
        [F, B-1]) is stored for later float predictions; default: the
        pre-binned edges 1..B-1.
        """
        if self.objective != "regression":
            raise ValueError("fit_binned supports objective='regression'")
        N, F = bin_indices.shape
        self._is_fitted = False
        self._trees_trained = 0
        self.forest = [{} for _ in range(self.n_estimators)]
        self.training_loss, self.eval_loss = [], []
        self._round_tables = []
        self._invalidate_prediction_cache()
        self.num_samples, self.num_features = N, F
        self.num_outputs = 1
        if era_id is None:  # one era: no host pass over the rows
            self._era_np = np.zeros(N, np.int32)
            self.unique_eras = torch.arange(1, device=self.device)
            self.era_indices = torch.zeros(N, dtype=torch.int32, device=self.device)
        else:
            self._era_np = _era_inverse(era_id)[1].astype(np.int32).reshape(-1)
            self.unique_eras = torch.arange(int(self._era_np.max()) + 1, device=self.device)
            self.era_indices = torch.from_numpy(self._era_np).to(self.device)
        self.num_eras = len(self.unique_eras)
        self.bin_indices = bin_indices.contiguous()
        self.Y_gpu = y.to(torch.float32).contiguous()
        w = torch.ones(N, dtype=torch.float32, device=self.device) if sample_weight is None else sample_weight
        self.sample_weight_gpu = w.to(torch.float32).contiguous()
        self._const_weight = None
        self.bin_edges = (bin_edges if bin_edges is not None else
                          torch.arange(1, self.num_bins, dtype=torch.float32, device=self.device).repeat(F, 1))
        self._prepare_monotonic_constraints(F)
        self._feature_ids = torch.arange(F, dtype=torch.int32, device=self.device)
        self.feature_indices = self._feature_ids
        if self.num_eras > 1:
            self._root_order = torch.from_numpy(np.argsort(self._era_np, kind="stable").astype(np.int32)).to(self.device)
        else:
            self._root_order = None
        if era_id is None:
            self._era_count = torch.full((1,), N, dtype=torch.int32, device=self.device)
        else:
            self._era_count = torch.from_numpy(np.bincount(self._era_np, minlength=self.num_eras).astype(np.int32)).to(
                self.device)
        self.root_node_indices = (self._root_order if self._root_order is not None else
                                  torch.arange(N, dtype=torch.int32, device=self.device))
        self.base_prediction = float(self.Y_gpu.double().mean().item())
        self.gradients = torch.full((N,), np.float32(self.base_prediction), dtype=torch.float32, device=self.device)
        self._setup_eval(None, None, None, None, None, "mse", classification=False)
        with torch.no_grad():
            self.grow_forest()
        del self.bin_indices
        del self.Y_gpu
        return self

    def _setup_eval(self, X_eval, y_eval, sample_weight_eval, eval_every_n_trees, early_stopping_rounds,
                    eval_metric, classification):
        if X_eval is not None and y_eval is not None:
            self.bin_indices_eval = self.bin_inference_data(X_eval)
            if classification:
                ye = self.label_encoder.transform(y_eval)
                self.Y_gpu_eval = torch.from_numpy(np.asarray(ye)).to(torch.int32).to(self.device)
                self.eval_metric = eval_metric if eval_metric != "mse" else "logloss"
            else:
                self.Y_gpu_eval = torch.from_numpy(np.asarray(y_eval)).to(torch.float32).to(self.device)
                self.eval_metric = eval_metric
            self.sample_weight_eval_gpu = torch.from_numpy(
                np.ascontiguousarray(sample_weight_eval, np.float32)).to(self.device)
            self.eval_every_n_trees = eval_every_n_trees
            self.early_stopping_rounds = early_stopping_rounds
            self._eval_scores = None  # batched engine: raw eval scores kept on the device
        else:
            self._eval_scores = None
            self.bin_indices_eval = None
            self.Y_gpu_eval = None
            self.sample_weight_eval_gpu = None
            self.eval_every_n_trees = None
            self.early_stopping_rounds = None

    def _fit_regression(self, X, y, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                        eval_every_n_trees, early_stopping_rounds, eval_metric):
        """warpgbm/core.py:883-958."""
        self.num_outputs = 1
        self._setup_common(X, y, era_id, sample_weight)
        if self.warm_start and self._is_fitted and self._trees_trained > 0:
            self.gradients = torch.full((self.num_samples,), float(self.base_prediction), dtype=torch.float32,
                                        device=self.device)
            tt = self._get_cached_regression_tree_tensor()
            if tt.size(0) > 0:
                _nk.predict_forest(self.bin_indices, tt.contiguous(), self.learning_rate, self.gradients)
        else:
            # unweighted target mean (warpgbm/core.py:926), accumulated in float64
            self.base_prediction = float(np.mean(np.asarray(y, dtype=np.float64)))
            self.gradients = torch.full((self.num_samples,), np.float32(self.base_prediction),
                                        dtype=torch.float32, device=self.device)
        self._setup_eval(X_eval, y_eval, sample_weight_eval, eval_every_n_trees, early_stopping_rounds,
                         eval_metric, classification=False)
        with torch.no_grad():
            self.grow_forest()
        del self.bin_indices
        del self.Y_gpu
        return self

    def _fit_classification(self, X, y_enc, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                            eval_every_n_trees, early_stopping_rounds, eval_metric):
        """warpgbm/core.py:1131-1224."""
        self.num_outputs = None
        self._setup_common(X, np.asarray(y_enc, dtype=np.float32), era_id, sample_weight)
        self.Y_gpu = torch.from_numpy(np.asarray(y_enc)).to(torch.int32).to(self.device)
        K = self.num_classes
        self.gradients = torch.zeros((self.num_samples, K), dtype=torch.float32, device=self.device)
        counts = np.bincount(np.asarray(y_enc), minlength=K)
        for k in range(K):
            prior = np.float32(counts[k]) / np.float32(len(y_enc))
            if prior > 0:
                self.gradients[:, k] = float(np.log(np.float32(prior) + np.float32(1e-10)))
        if self.warm_start and self._is_fitted and self._trees_trained > 0:
            tts = self._get_cached_multiclass_tree_tensors()
            for k, tt in enumerate(tts):
                if tt is not None and tt.size(0) > 0:
                    col = self.gradients[:, k].contiguous()
                    _nk.predict_forest(self.bin_indices, tt.contiguous(), self.learning_rate, col)
                    self.gradients[:, k] = col
        self._setup_eval(X_eval, y_eval, sample_weight_eval, eval_every_n_trees, early_stopping_rounds,
                         eval_metric, classification=True)
        with torch.no_grad():
            self.grow_forest_multiclass()
        del self.bin_indices
        del self.Y_gpu
        return self

    def _fit_multi_regression(self, X, y, era_id, X_eval, y_eval, sample_weight, sample_weight_eval,
                              eval_every_n_trees, early_stopping_rounds, eval_metric):
        """Vector-leaf multi-output regression (warpgbm/core.py:960-1129).

        Targets stay as given on the device: a non-finite target is a missing
        one (weight 0 in the gradients, the base and the losses), which is what
        the reference's NaN fill plus finite mask amounts to."""
        M = y.shape[1]
        if self.warm_start and self._is_fitted and self.num_outputs is not None and self.num_outputs != M:
            raise ValueError("Number of target outputs must match the fitted model when warm_start=True. "
                             f"Got {M} and {self.num_outputs}.")
        self.num_outputs = M
        y32 = np.ascontiguousarray(y, dtype=np.float32)
        w2 = self._expand_multi_output_weight(np.asarray(sample_weight, np.float32), M)
        self._setup_common(X, y32, era_id, w2)
        if self.warm_start and self._is_fitted and self._trees_trained > 0:
            base = torch.as_tensor(self.base_prediction, dtype=torch.float32).reshape(-1)
            if base.numel() != M:
                raise ValueError("Stored base_prediction is incompatible with the requested number "
                                 f"of outputs ({M}).")
            self.base_prediction = base.to(self.device)
            self.gradients = self._multi_scores(self.bin_indices)
        else:
            # weighted mean of the valid targets per output, 0 where an output has no weight
            # (core.py:1055-1063), accumulated in float64
            valid = np.isfinite(y32)
            ew = np.where(valid, w2, 0).astype(np.float64)
            num = (np.where(valid, y32, 0).astype(np.float64) * ew).sum(axis=0)
            den = ew.sum(axis=0)
            base = np.where(den > 0, num / np.maximum(den, 1e-12), 0.0).astype(np.float32)
            self.base_prediction = torch.from_numpy(base).to(self.device)
            self.gradients = self.base_prediction.reshape(1, M).repeat(self.num_samples, 1).contiguous()
        if X_eval is not None and y_eval is not None:
            self.bin_indices_eval = self.bin_inference_data(X_eval)
            self.Y_gpu_eval = torch.from_numpy(np.ascontiguousarray(y_eval, np.float32)).to(self.device)
            self.sample_weight_eval_gpu = torch.from_numpy(np.ascontiguousarray(
                self._expand_multi_output_weight(np.asarray(sample_weight_eval, np.float32), M))).to(self.device)
            self.eval_every_n_trees = eval_every_n_trees
            self.early_stopping_rounds = early_stopping_rounds
            self.eval_metric = eval_metric
            self._eval_scores = None
        else:
            self._setup_eval(None, None, None, None, None, eval_metric, classification=False)
        with torch.no_grad():
            self.grow_forest_multi_regression()
        del self.bin_indices
        del self.Y_gpu
        return self

    @staticmethod
    def _expand_multi_output_weight(w, n_outputs):
        """1-D weights repeated over the outputs (warpgbm/core.py:580-586)."""
        if w.ndim == 1:
            return np.repeat(w[:, None], n_outputs, axis=1)
        if w.ndim == 2:
            return w
        raise ValueError(f"sample weights must be 1D or 2D, got shape {w.shape}")

    # ------------------------------------------------------------ binning
    def preprocess_gpu_data(self, X_np, Y_np, era_id_np):
        """Codes, eras and edges on the device (warpgbm/core.py:1226-1285)."""
        with torch.no_grad():
            N, F = X_np.shape
            self.num_samples, self.num_features = N, F
            Y_gpu = torch.from_numpy(np.ascontiguousarray(Y_np)).to(dtype=torch.float32, device=self.device)
            uniq, inv = _era_inverse(era_id_np)
            self._era_np = inv.astype(np.int32).reshape(-1)
            era_indices = torch.from_numpy(self._era_np).to(self.device)
            unique_eras = torch.from_numpy(uniq).to(self.device)
            if np.issubdtype(X_np.dtype, np.integer) and np.all(X_np.max(axis=0) < self.num_bins):
                print("Detected pre-binned integer input — skipping quantile binning.")
                bin_indices = _to_device(X_np, self.device, None).to(torch.int8).contiguous()
                bin_edges = torch.arange(1, self.num_bins, dtype=torch.float32).repeat(F, 1).to(self.device)
                return bin_indices, era_indices, bin_edges, unique_eras, Y_gpu
            X_float = _to_device(X_np, self.device)
            bin_edges = self._quantile_edges(X_float)
            bin_indices = self._bin_matrix(X_float, bin_edges)
            del X_float
            return bin_indices, era_indices, bin_edges, unique_eras, Y_gpu

    def _quantile_edges(self, X_float):
        """Per-feature float32 quantile edges (torch.quantile semantics,
        warpgbm/core.py:1271-1278) by radix selection on the device, no sort.
        Columns go in blocks whose column-major copy fits 2 GiB."""
        N, F = X_float.shape
        B = self.num_bins
        X_float = X_float.contiguous()
        edges = torch.empty((F, B - 1), dtype=torch.float32, device=self.device)
        blk = max(1, min(F, (2 << 30) // max(4 * N, 1)))
        ws = self._ws.get("quantile", _lib.query("wgbm_quantile_edges_workspace_bytes", N, blk), self.device)
        for f0 in range(0, F, blk):
            f1 = min(F, f0 + blk)
            _lib.call("wgbm_quantile_edges", X_float.data_ptr() + 4 * f0, N, f1 - f0, F, B, edges[f0].data_ptr(),
                      B - 1, ws.data_ptr(), ws.numel(), _stream(self.device))
        self._ws._bufs.pop("quantile", None)  # the column-major copy is not needed after binning
        return edges

    def _bin_matrix(self, X_float, edges):
        N, F = X_float.shape
        out = torch.empty((N, F), dtype=torch.int8, device=self.device)
        ne = edges.shape[1]
        ws = self._ws.get("bin", _lib.query("wgbm_bin_matrix_workspace_bytes", F, ne), self.device)
        _lib.call("wgbm_bin_matrix", X_float.data_ptr(), N, F, edges.contiguous().data_ptr(), ne, out.data_ptr(),
                  ws.data_ptr(), ws.numel(), _stream(self.device))
        return out

    def bin_data_with_existing_edges(self, X_np):
        """Inference binning with stored edges, one launch (core.py:1882-1900)."""
        X_float = _to_device(X_np, self.device)
        edges = self.bin_edges.to(self.device).contiguous()
        return self._bin_matrix(X_float.contiguous(), edges)

    def bin_inference_data(self, X_np):
        """Pre-binned bypass or binning with stored edges (core.py:1964-1983)."""
        pre = False
        if np.issubdtype(X_np.dtype, np.integer) and X_np.shape[1] == self.num_features:
            if np.all(X_np.max(axis=0) < self.num_bins):
                print("Detected pre-binned input at predict-time — skipping binning.")
                pre = True
        if pre:
            return _to_device(X_np, self.device, None).to(torch.int8).contiguous()
        return self.bin_data_with_existing_edges(X_np)

    # Share of the device's free memory the batched engine's workspace may
    # take; above it the per-node engine (which needs about one node's
    # histograms at a time) trains instead.
    _BATCHED_MEM_SHARE = 0.8
    # Bytes of node tables (device and host) one wgbm_fit_rounds call may
    # write; more rounds go in several calls.
    _TABLE_BYTES_PER_CALL = 1 << 30

    def _n_channels(self):
        if self.objective in ("multiclass", "binary"):
            return self.num_classes
        return self.num_outputs if self.objective == "multi_regression" else 1

    def _use_plain(self):
        """The per-node engine runs when asked for, for vector-leaf
        multi-output regression, and for fits the batched engine cannot
        hold: trees deeper than its level tables (depth above 20, or
        classes x 2^depth of 2^24 or more), more than 256 classes, classes x
        rows of 2^31 or more, or a workspace above a share of the device's
        free memory.  Both engines grow identical trees."""
        if self.engine == "plain" or self.objective == "multi_regression":
            return True
        K = self._n_channels()
        if self.max_depth > 20 or (K << self.max_depth) >= (1 << 24) or K > 256:
            return True
        n = getattr(self, "num_samples", None)
        if n is None:
            return False
        if n * K >= 0x7fffffff:
            return True
        need = _lib.query("wgbm_fit_workspace_bytes", ctypes.byref(self._fit_config(self._n_sampled(), K)))
        if need == 0:
            return True
        dev = torch.device(self.device)
        if dev.type == "cuda" and torch.cuda.is_available():
            free, _ = torch.cuda.mem_get_info(dev)
            # the device's shared fit workspace is reused (or replaced) by this fit
            idx = dev.index if dev.index is not None else torch.cuda.current_device()
            free += max((b.numel() for (i, _), b in _FIT_WORKSPACES.items() if i == idx), default=0)
            if need > self._BATCHED_MEM_SHARE * free:
                print(f"Batched engine needs {need / 2**30:.1f} GiB of workspace "
                      f"({free / 2**30:.1f} GiB free): training per node instead.")
                return True
        return False

    def _rounds_per_call(self, K):
        """Rounds whose node tables fit _TABLE_BYTES_PER_CALL."""
        per_round = K * (2 ** (self.max_depth + 1) - 1) * (20 + 4 * self.num_eras)
        return max(1, int(self._TABLE_BYTES_PER_CALL // per_round))

    # ------------------------------------------------------------ plain per-node engine
    def compute_histograms(self, sample_indices, feature_indices, grad, hess):
        """G, H [E, F, B] of one node (warpgbm/core.py:1287-1308)."""
        G = torch.zeros((self.num_eras, len(feature_indices), self.num_bins), device=self.device,
                        dtype=torch.float32)
        H = torch.zeros_like(G)
        _nk.compute_histogram3(self.bin_indices, grad, hess, sample_indices, feature_indices, self.era_indices,
                               G, H, self.num_bins, self.threads_per_block, self.rows_per_thread)
        return G, H

    compute_histograms_multiclass = compute_histograms

    def find_best_split(self, gradient_histogram, hessian_histogram):
        """Per-era scan plus DES/monotone selection of one node (core.py:1310-1362)."""
        _nk.compute_split(gradient_histogram, hessian_histogram, self.min_split_gain, self.min_child_weight,
                          self.L2_reg, self.per_era_gain, self.per_era_direction, self.threads_per_block)
        out = self._select_out
        _lib.call("wgbm_select_split", self.per_era_gain.data_ptr(), self.per_era_direction.data_ptr(),
                  self.num_eras, self.per_era_gain.shape[1], self.num_bins, float(self.min_split_gain),
                  self.monotonic_required_split_direction_gpu_.data_ptr(), self.feat_indices_tree.data_ptr(),
                  out.data_ptr(), _stream(self.device))
        f, b = out.cpu().tolist()
        return int(f), int(b)

    def _tree_scale(self, grad, hess, out=None):
        """Statistics of a tree's root rows (device int32[4]): the scale of
        its leaves' exact sums, as the batched engine scales them."""
        if out is None:
            out = torch.empty(4, dtype=torch.int32, device=self.device)
        ws = self._ws.get("row_stats", _lib.query("wgbm_row_stats_workspace_bytes"), self.device)
        rows = self.root_node_indices
        _lib.call("wgbm_row_stats", grad.data_ptr(), hess.data_ptr(), rows.data_ptr(), rows.numel(),
                  self.num_samples, out.data_ptr(), ws.data_ptr(), ws.numel(), _stream(self.device))
        return out

    def _plain_leaf(self, node_indices, class_k):
        n = node_indices.numel()
        ws = self._ws.get("leaf", _lib.query("wgbm_leaf_workspace_bytes", n), self.device)
        if self.objective == "multi_regression" and class_k is None:
            # vector leaf: per output m, sum(g_m)/(sum(h_m)+lambda) over the node's
            # rows, and that output's score column updated (core.py:1366-1375)
            M = self.num_outputs
            for m in range(M):
                _lib.call("wgbm_leaf_value", self.grad_out[m].data_ptr(), self.hess_out[m].data_ptr(),
                          node_indices.data_ptr(), n, self.num_samples, float(self.L2_reg),
                          float(self.learning_rate), self.gradients.data_ptr(), M, m,
                          self._leaf_out[m].data_ptr(), self._scale_out[m].data_ptr(), ws.data_ptr(), ws.numel(),
                          _stream(self.device))
            return {"leaf_value": self._leaf_out[:M].cpu().tolist(), "samples": int(n)}
        K = self.num_classes if class_k is not None else 1
        _lib.call("wgbm_leaf_value", self.grad.data_ptr(), self.hess.data_ptr(), node_indices.data_ptr(), n,
                  self.num_samples, float(self.L2_reg), float(self.learning_rate), self.gradients.data_ptr(), K,
                  class_k or 0, self._leaf_out.data_ptr(), self._scale.data_ptr(), ws.data_ptr(), ws.numel(),
                  _stream(self.device))
        return {"leaf_value": float(self._leaf_out.item()), "samples": int(n)}

    def grow_tree(self, gradient_histogram, hessian_histogram, node_indices, depth, class_k=None):
        """Depth-first growth with one launch set per node (warpgbm/core.py:1364-1470)."""
        if depth == self.max_depth:
            return self._plain_leaf(node_indices, class_k)
        local_feature, best_bin = self.find_best_split(gradient_histogram, hessian_histogram)
        if local_feature == -1:
            return self._plain_leaf(node_indices, class_k)
        gfeat = int(self._feat_tree_host[local_feature])
        self.per_era_feature_importance_[:, gfeat] += self.per_era_gain[:, local_feature, best_bin].cpu().numpy()
        n = node_indices.numel()
        out = torch.empty(n, dtype=torch.int32, device=self.device)
        ws = self._ws.get("part", _lib.query("wgbm_partition_workspace_bytes", n), self.device)
        _lib.call("wgbm_partition_rows", self.bin_indices.data_ptr(), self.num_samples, self.num_features,
                  node_indices.data_ptr(), n, gfeat, best_bin, out.data_ptr(), self._nleft.data_ptr(),
                  ws.data_ptr(), ws.numel(), _stream(self.device))
        nl = int(self._nleft.item())
        left_indices, right_indices = out[:nl], out[nl:]
        if nl <= n - nl:
            gl, hl = self.compute_histograms(left_indices, self.feat_indices_tree, self.grad, self.hess)
            gr, hr = torch.empty_like(gl), torch.empty_like(hl)
            _lib.call("wgbm_hist_subtract", gradient_histogram.data_ptr(), hessian_histogram.data_ptr(),
                      gl.data_ptr(), hl.data_ptr(), gr.data_ptr(), hr.data_ptr(), gl.numel(), _stream(self.device))
        else:
            gr, hr = self.compute_histograms(right_indices, self.feat_indices_tree, self.grad, self.hess)
            gl, hl = torch.empty_like(gr), torch.empty_like(hr)
            _lib.call("wgbm_hist_subtract", gradient_histogram.data_ptr(), hessian_histogram.data_ptr(),
                      gr.data_ptr(), hr.data_ptr(), gl.data_ptr(), hl.data_ptr(), gr.numel(), _stream(self.device))
        left = self.grow_tree(gl, hl, left_indices, depth + 1, class_k)
        right = self.grow_tree(gr, hr, right_indices, depth + 1, class_k)
        return {"feature": self._feature_ids[gfeat], "bin": int(best_bin), "left": left, "right": right}

    def _plain_setup(self, k):
        self.per_era_gain = torch.zeros(self.num_eras, k, self.num_bins - 1, device=self.device, dtype=torch.float32)
        self.per_era_direction = torch.zeros_like(self.per_era_gain)
        self._select_out = torch.empty(2, dtype=torch.int32, device=self.device)
        self._nleft = torch.empty(1, dtype=torch.int32, device=self.device)
        self._leaf_out = torch.empty(1, dtype=torch.float32, device=self.device)

    def _tree_features(self, k):
        if self.colsample_bytree < 1.0:
            self.feat_indices_tree = torch.randperm(self.num_features, device=self.device,
                                                    dtype=torch.int32)[:k].contiguous()
        else:
            self.feat_indices_tree = self.feature_indices
        self._feat_tree_host = self.feat_indices_tree.cpu().numpy()

    # ------------------------------------------------------------ forests
    def _n_sampled(self):
        if self.colsample_bytree < 1.0:
            return max(1, int(self.colsample_bytree * self.num_features))
        return self.num_features

    def grow_forest(self):
        """Regression boosting loop (warpgbm/core.py:1518-1580)."""
        if not self.warm_start or not self._is_fitted or self.per_era_feature_importance_ is None:
            self.training_loss = []
            self.eval_loss = []
            self.per_era_feature_importance_ = np.zeros((self.num_eras, self.num_features), dtype=np.float32)
        self.stop = False
        k = self._n_sampled()
        start = self._trees_trained if self.warm_start and self._is_fitted else 0
        if len(self.forest) < self.n_estimators:
            self.forest.extend([{} for _ in range(self.n_estimators - len(self.forest))])
        if self._use_plain():
            self._round_tables = None  # trees without level tables: the heap forest is built from the dicts
            self._plain_setup(k)
            self.grad = torch.empty(self.num_samples, dtype=torch.float32, device=self.device)
            self.hess = torch.empty_like(self.grad)
            for i in range(start, self.n_estimators):
                _lib.call("wgbm_grad_regression", self.Y_gpu.data_ptr(), self.gradients.data_ptr(),
                          self.sample_weight_gpu.data_ptr(), self.num_samples, self.grad.data_ptr(),
                          self.hess.data_ptr(), _stream(self.device))
                self._tree_features(k)
                self._scale = self._tree_scale(self.grad, self.hess)
                G, H = self.compute_histograms(self.root_node_indices, self.feat_indices_tree, self.grad, self.hess)
                self.forest[i] = self.grow_tree(G, H, self.root_node_indices, 0)
                self._trees_trained = i + 1
                self._invalidate_prediction_cache()
                self.compute_eval(i)
                if self.stop:
                    break
        else:
            self._batched_forest(start, k, n_classes=1)
        self.feature_importance_ = self.per_era_feature_importance_.sum(axis=0)
        self._is_fitted = True
        print(f"Finished training forest. Total trees: {self._trees_trained}")

    def grow_forest_multiclass(self):
        """K trees per round (warpgbm/core.py:1719-1791)."""
        if not self.warm_start or not self._is_fitted or self.per_era_feature_importance_ is None:
            self.training_loss = []
            self.eval_loss = []
            self.per_era_feature_importance_ = np.zeros((self.num_eras, self.num_features), dtype=np.float32)
            self.forest = []
        self.stop = False
        k = self._n_sampled()
        start = self._trees_trained if self.warm_start and self._is_fitted else 0
        K = self.num_classes
        if self._use_plain():
            self._round_tables = None
            self._plain_setup(k)
            g_all = torch.empty((K, self.num_samples), dtype=torch.float32, device=self.device)
            h_all = torch.empty_like(g_all)
            for i in range(start, self.n_estimators):
                _lib.call("wgbm_grad_softmax", self.gradients.data_ptr(), self.Y_gpu.data_ptr(),
                          self.sample_weight_gpu.data_ptr(), self.num_samples, K, g_all.data_ptr(),
                          h_all.data_ptr(), _stream(self.device))
                self._tree_features(k)
                trees_k = []
                for c in range(K):
                    self.grad = g_all[c]
                    self.hess = h_all[c]
                    self._scale = self._tree_scale(self.grad, self.hess)
                    G, H = self.compute_histograms(self.root_node_indices, self.feat_indices_tree, self.grad,
                                                   self.hess)
                    trees_k.append(self.grow_tree(G, H, self.root_node_indices, 0, class_k=c))
                self.forest.append(trees_k)
                self._trees_trained = i + 1
                self._invalidate_prediction_cache()
                self.compute_eval_multiclass(i)
                if self.stop:
                    break
        else:
            self._batched_forest(start, k, n_classes=K)
        self.feature_importance_ = self.per_era_feature_importance_.sum(axis=0)
        self._is_fitted = True
        print(f"Finished training multiclass forest. Total rounds: {self._trees_trained} "
              f"({self._trees_trained * self.num_classes} trees)")

    def grow_forest_multi_regression(self):
        """One tree per round with splits shared by all outputs and a vector
        of leaf values (warpgbm/core.py:1619-1717): the split search runs on
        the per-row sums of the outputs' gradients and hessians, the leaves
        on each output's own."""
        if not self.warm_start or not self._is_fitted or self.per_era_feature_importance_ is None:
            self.training_loss = []
            self.eval_loss = []
            self.per_era_feature_importance_ = np.zeros((self.num_eras, self.num_features), dtype=np.float32)
        self.stop = False
        k = self._n_sampled()
        start = self._trees_trained if self.warm_start and self._is_fitted else 0
        if len(self.forest) < self.n_estimators:
            self.forest.extend([{} for _ in range(self.n_estimators - len(self.forest))])
        self._round_tables = None
        self._plain_setup(k)
        N, M = self.num_samples, self.num_outputs
        self._leaf_out = torch.empty(M, dtype=torch.float32, device=self.device)
        self.grad_out = torch.empty((M, N), dtype=torch.float32, device=self.device)
        self.hess_out = torch.empty_like(self.grad_out)
        self.grad = torch.empty(N, dtype=torch.float32, device=self.device)
        self.hess = torch.empty_like(self.grad)
        self._scale_out = torch.empty((M, 4), dtype=torch.int32, device=self.device)
        for i in range(start, self.n_estimators):
            _lib.call("wgbm_grad_multi_regression", self.Y_gpu.data_ptr(), self.gradients.data_ptr(),
                      self.sample_weight_gpu.data_ptr(), N, M, self.grad_out.data_ptr(), self.hess_out.data_ptr(),
                      self.grad.data_ptr(), self.hess.data_ptr(), _stream(self.device))
            self._tree_features(k)
            for m in range(M):  # each output's leaves are scaled by that output's root rows
                self._tree_scale(self.grad_out[m], self.hess_out[m], out=self._scale_out[m])
            G, H = self.compute_histograms(self.root_node_indices, self.feat_indices_tree, self.grad, self.hess)
            self.forest[i] = self.grow_tree(G, H, self.root_node_indices, 0)
            self._trees_trained = i + 1
            self._invalidate_prediction_cache()
            self.compute_eval_multi_regression(i)
            if self.stop:
                break
        del self.grad_out, self.hess_out
        self.feature_importance_ = self.per_era_feature_importance_.sum(axis=0)
        self._is_fitted = True
        print(f"Finished training multi-output forest. Total trees: {self._trees_trained}")

    # ------------------------------------------------------------ batched engine
    def _fit_config(self, k, n_classes):
        c = _lib.FitConfig()
        c.n_rows = self.num_samples
        c.n_cols = self.num_features
        c.n_feat = k
        c.n_bins = self.num_bins
        c.n_eras = self.num_eras
        c.n_classes = n_classes
        c.max_depth = self.max_depth
        c.objective = 0 if self.objective == "regression" else 1
        c.learning_rate = float(self.learning_rate)
        c.min_split_gain = float(self.min_split_gain)
        c.min_child_weight = float(self.min_child_weight)
        c.lambda_ = float(self.L2_reg)
        # Equal weights make regression hessians constant: the histogram
        # kernel then keeps one LDS plane (checked once per fit on device).
        w = getattr(self, "sample_weight_gpu", None)
        if getattr(self, "_const_weight", None) is None:
            if w is None:
                # a sizing query before the weights exist: assume equal weights
                c.const_weight = 1
                return c
            self._const_weight = bool(w.numel() == 0 or bool((w == w[0]).all()))
        c.const_weight = int(self._const_weight)
        return c

    def _batched_forest(self, start, k, n_classes):
        """Run wgbm_fit_rounds; without an eval set all rounds go in one call
        (no host sync until the end), with one they go round by round."""
        total = self.n_estimators - start
        if total <= 0:
            return
        cfg = self._fit_config(k, n_classes)
        if self.autotune:
            from . import autotune
            self.hist_config_ = autotune.tune_fit(self, cfg, k)
        ws = _fit_workspace(self.device, _lib.query("wgbm_fit_workspace_bytes", ctypes.byref(cfg)))
        every = self.eval_every_n_trees
        if every is not None:
            # raw scores of the eval rows under the trees so far (warm start
            # included); each call adds its new trees to them
            if n_classes == 1 and self.objective == "regression":
                self._eval_scores = self.predict_binned(self.bin_indices_eval).reshape(-1, 1)
            else:
                self._eval_scores = self._raw_multiclass_scores(self.bin_indices_eval)
        i = start
        cap = self._rounds_per_call(n_classes)
        while i < self.n_estimators and not self.stop:
            if every is None:
                n = self.n_estimators - i
            else:
                # end the call at the next tree the eval runs after (tree index
                # a multiple of eval_every_n_trees): rounds in between only log
                # their training loss, computed on the device
                n = min(-(-i // every) * every - i + 1, self.n_estimators - i)
            n = min(n, cap)
            feats = None
            if self.colsample_bytree < 1.0:
                feats = torch.stack([torch.randperm(self.num_features, device=self.device, dtype=torch.int32)[:k]
                                     for _ in range(n)]).contiguous()
            tables = self._run_rounds(cfg, ws, n, feats, with_loss=every is not None)
            self._append_rounds(tables, n, n_classes, feats)
            self._trees_trained = i + n
            self._invalidate_prediction_cache()
            if every is not None:
                self._eval_add_rounds(tables, n, n_classes)
                for r in range(n):
                    loss = float(tables["train_loss"][r])
                    if n_classes == 1 and self.objective == "regression":
                        self.compute_eval(i + r, train_loss=loss)
                    else:
                        self.compute_eval_multiclass(i + r, train_loss=loss)
            i += n

    def _run_rounds(self, cfg, ws, n, feats, with_loss=False, read_back=True):
        """n boosting rounds in one wgbm_fit_rounds call; their node tables
        (host arrays), or None without read_back."""
        K = cfg.n_classes
        NH = 2 ** (self.max_depth + 1) - 1
        dev = self.device
        shape = (n, K, NH)
        # All tables in one device buffer, read back with one copy.
        m = n * K * NH
        E = self.num_eras
        buf = torch.empty(m * (5 + E) + (n if with_loss else 0), dtype=torch.int32, device=dev)
        t = {key: buf[i * m:(i + 1) * m].view(shape) for i, key in enumerate(("status", "feat", "bin", "count"))}
        t["value"] = buf[4 * m:5 * m].view(torch.float32).view(shape)
        t["era_gain"] = buf[5 * m:(5 + E) * m].view(torch.float32).view(shape + (E,))
        if with_loss:
            t["train_loss"] = buf[(5 + E) * m:].view(torch.float32)
        io = _lib.FitIO()
        io.codes = self.bin_indices.data_ptr()
        io.y = self.Y_gpu.data_ptr() if cfg.objective == 0 else None
        io.labels = self.Y_gpu.data_ptr() if cfg.objective == 1 else None
        io.weight = self.sample_weight_gpu.data_ptr()
        io.root_order = self._root_order.data_ptr() if self._root_order is not None else None
        io.era_count = self._era_count.data_ptr()
        io.mono_required = self.monotonic_required_split_direction_gpu_.data_ptr()
        io.feat_idx = feats.data_ptr() if feats is not None else None
        io.scores = self.gradients.data_ptr()
        io.node_status = t["status"].data_ptr()
        io.node_feat = t["feat"].data_ptr()
        io.node_bin = t["bin"].data_ptr()
        io.node_value = t["value"].data_ptr()
        io.node_count = t["count"].data_ptr()
        io.node_era_gain = t["era_gain"].data_ptr()
        io.train_loss = t["train_loss"].data_ptr() if with_loss else None
        _lib.call("wgbm_fit_rounds", ctypes.byref(cfg), ctypes.byref(io), n, ws.data_ptr(), ws.numel(),
                  _stream(dev))
        if not read_back:
            return None
        host = buf.cpu().numpy()
        tables = {"status": host[:m].reshape(shape), "feat": host[m:2 * m].reshape(shape),
                  "bin": host[2 * m:3 * m].reshape(shape), "count": host[3 * m:4 * m].reshape(shape),
                  "value": host[4 * m:5 * m].view(np.float32).reshape(shape),
                  "era_gain": host[5 * m:(5 + E) * m].view(np.float32).reshape(shape + (E,))}
        if with_loss:
            tables["train_loss"] = host[(5 + E) * m:].view(np.float32)
        self._last_tables = tables
        return tables

    def _advance_rounds(self, cfg, n_rounds, feats_gen=None):
        """Boost n_rounds on the current scores, discarding the trees: the
        state from which the autotuner times its candidates (a fit's trees a
        few rounds in, not its first)."""
        ws = _fit_workspace(self.device, _lib.query("wgbm_fit_workspace_bytes", ctypes.byref(cfg)))
        feats = None
        if cfg.n_feat < self.num_features:
            feats = torch.stack([torch.randperm(self.num_features, generator=feats_gen, device=self.device,
                                                dtype=torch.int32)[:cfg.n_feat] for _ in range(n_rounds)])
            feats = feats.contiguous()
        self._run_rounds(cfg, ws, n_rounds, feats)

    def _trial_rounds_ms(self, cfg, n_rounds, feats_gen=None):
        """Device time (ms) of one wgbm_fit_rounds call of n_rounds boosting
        rounds on a copy of the scores, between events on the fit's stream:
        the model is left as it was."""
        ws = _fit_workspace(self.device, _lib.query("wgbm_fit_workspace_bytes", ctypes.byref(cfg)))
        feats = None
        if cfg.n_feat < self.num_features:
            feats = torch.stack([torch.randperm(self.num_features, generator=feats_gen, device=self.device,
                                                dtype=torch.int32)[:cfg.n_feat] for _ in range(n_rounds)])
            feats = feats.contiguous()
        saved, last = self.gradients, getattr(self, "_last_tables", None)
        self.gradients = saved.clone()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        try:
            e0.record()
            self._run_rounds(cfg, ws, n_rounds, feats, read_back=False)
            e1.record()
            e1.synchronize()
        finally:
            self.gradients, self._last_tables = saved, last
        return e0.elapsed_time(e1)

    def _eval_add_rounds(self, tables, n, K):
        """Add the n new rounds' trees to the device eval scores: the rounds'
        node tables compile to a heap forest (no per-tree host flattening) whose
        traversal gives each eval row's sum over those trees."""
        from .serving import HeapForest
        NH = tables["status"].shape[-1]
        hf = HeapForest.from_tables(tables["status"][:n].reshape(n * K, NH), tables["feat"][:n].reshape(n * K, NH),
                                    tables["bin"][:n].reshape(n * K, NH), tables["value"][:n].reshape(n * K, NH),
                                    self.max_depth, self.learning_rate, self.num_features, [0.0] * K,
                                    n_classes=K, device=self.device)
        self._eval_scores += hf.predict(self.bin_indices_eval)

    def _append_rounds(self, tables, n, K, feats):
        """Node tables of n rounds -> the reference's nested tree dicts, and
        per-era importances (gain of every split added to its feature).
        Heap slot status: 1 split, 2 leaf, 0 not in the tree."""
        st, ft, bt, vt, ct, eg = (tables[k] for k in ("status", "feat", "bin", "value", "count", "era_gain"))
        NH = st.shape[-1]
        if getattr(self, "_round_tables", None) is not None:
            self._round_tables.append({"status": st[:n], "feat": ft[:n], "bin": bt[:n], "value": vt[:n]})
        pre = _heap_preorder(NH)
        # gains added split by split in growth (depth-first) order, as the
        # per-node path does: rounds, then classes, then splits in preorder
        rr, cc, pp = np.nonzero(st[:n][:, :, pre] == 1)
        if len(rr):
            hh = pre[pp]
            np.add.at(self.per_era_feature_importance_.T, ft[rr, cc, hh], eg[rr, cc, hh])
        ids = self._feature_ids
        # Slots of all n*K trees by flat index t = (r*K + c)*NH + h: the
        # leaves first, then the splits from the last slot back, so a split's
        # children (t + h + 1, t + h + 2) are built before it.
        flat = st[:n].reshape(-1)
        nodes = [None] * flat.size
        lf = np.flatnonzero(flat == 2)
        for t, v, c in zip(lf.tolist(), vt[:n].reshape(-1)[lf].tolist(), ct[:n].reshape(-1)[lf].tolist()):
            nodes[t] = {"leaf_value": v, "samples": c}
        sp = np.flatnonzero(flat == 1)[::-1]
        fs = ft[:n].reshape(-1)[sp]
        views = {f: ids[f] for f in np.unique(fs).tolist()}  # 0-d views into the feature-id vector
        for t, l, f, b in zip(sp.tolist(), (sp + sp % NH + 1).tolist(), fs.tolist(), bt[:n].reshape(-1)[sp].tolist()):
            nodes[t] = {"feature": views[f], "bin": b, "left": nodes[l], "right": nodes[l + 1]}
        for r in range(n):
            trees = [nodes[(r * K + c) * NH] for c in range(K)]
            if self.objective == "regression":
                self.forest[self._trees_trained + r] = trees[0]
            else:
                self.forest.append(trees)

    # ------------------------------------------------------------ evaluation
    def get_eval_metric(self, y_true, y_pred, sample_weight=None):
        if self.eval_metric == "mse":
            return weighted_mse_torch(y_true, y_pred, sample_weight=sample_weight).item()
        if self.eval_metric == "corr":
            return weighted_corr_loss_torch(y_true, y_pred, sample_weight=sample_weight).item()
        if self.eval_metric == "rmsle":
            return weighted_rmsle_torch(y_true, y_pred, sample_weight=sample_weight).item()
        raise ValueError(f"Invalid eval_metric: {self.eval_metric}.")

    def compute_eval(self, i, train_loss=None):
        """Training loss every tree, eval loss every n trees, early stop (core.py:1488-1516).
        train_loss: the loss after tree i when the caller already has it (the
        batched engine computes it on the device)."""
        if self.eval_every_n_trees is None:
            return
        if train_loss is None:
            train_loss = weighted_mse_torch(self.Y_gpu, self.gradients, sample_weight=self.sample_weight_gpu).item()
        self.training_loss.append(train_loss)
        if i % self.eval_every_n_trees == 0:
            preds = (self._eval_scores[:, 0] if getattr(self, "_eval_scores", None) is not None
                     else self.predict_binned(self.bin_indices_eval))
            loss = self.get_eval_metric(self.Y_gpu_eval, preds, sample_weight=self.sample_weight_eval_gpu)
            self.eval_loss.append(loss)
            if len(self.eval_loss) > self.early_stopping_rounds:
                if self.eval_loss[-(self.early_stopping_rounds + 1)] < self.eval_loss[-1]:
                    self.stop = True
            print(f"🌲 Tree {i+1}/{self.n_estimators} | Train MSE: {train_loss:.6f} | "
                  f"Eval {self.eval_metric}: {loss:.6f}")

    def compute_eval_multiclass(self, i, train_loss=None):
        """warpgbm/core.py:1793-1845 (train_loss: as in compute_eval)."""
        if self.eval_every_n_trees is None:
            return
        if train_loss is None:
            probs = softmax(self.gradients, dim=1)
            train_loss = weighted_log_loss_torch(self.Y_gpu, probs, sample_weight=self.sample_weight_gpu).item()
        self.training_loss.append(train_loss)
        if i % self.eval_every_n_trees == 0:
            ep = (softmax(self._eval_scores, dim=1) if getattr(self, "_eval_scores", None) is not None
                  else self.predict_proba_binned(self.bin_indices_eval))
            if self.eval_metric == "accuracy":
                loss = 1.0 - weighted_accuracy_torch(self.Y_gpu_eval, torch.argmax(ep, dim=1),
                                                     sample_weight=self.sample_weight_eval_gpu).item()
            else:
                loss = weighted_log_loss_torch(self.Y_gpu_eval, ep, sample_weight=self.sample_weight_eval_gpu).item()
            self.eval_loss.append(loss)
            if len(self.eval_loss) > self.early_stopping_rounds:
                if self.eval_loss[-(self.early_stopping_rounds + 1)] < self.eval_loss[-1]:
                    self.stop = True
            print(f"🌲 Round {i+1}/{self.n_estimators} ({self.num_classes} trees) | "
                  f"Train LogLoss: {train_loss:.6f} | Eval {self.eval_metric}: {loss:.6f}")

    def compute_eval_multi_regression(self, i):
        """warpgbm/core.py:1582-1617."""
        if self.eval_every_n_trees is None:
            return
        train_loss = weighted_multi_mse_torch(self.Y_gpu, self.gradients, sample_weight=self.sample_weight_gpu).item()
        self.training_loss.append(train_loss)
        if i % self.eval_every_n_trees == 0:
            preds = self.predict_binned(self.bin_indices_eval)
            loss = weighted_multi_mse_torch(self.Y_gpu_eval, preds, sample_weight=self.sample_weight_eval_gpu).item()
            self.eval_loss.append(loss)
            if len(self.eval_loss) > self.early_stopping_rounds:
                if self.eval_loss[-(self.early_stopping_rounds + 1)] < self.eval_loss[-1]:
                    self.stop = True
            print(f"🌲 Tree {i+1}/{self.n_estimators} | Train multi_regression MSE: {train_loss:.6f} | "
                  f"Eval MSE: {loss:.6f}")

    # ------------------------------------------------------------ prediction
    def _base_vector(self):
        """Per-channel starting scores of a prediction."""
        if self.objective == "regression":
            return [float(self.base_prediction)]
        if self.objective == "multi_regression":
            return torch.as_tensor(self.base_prediction, dtype=torch.float32).reshape(-1).cpu().tolist()
        K = self.num_classes
        return [float(np.log(np.float32(1.0 / K) + 1e-10))] * K  # the predict-time prior (core.py:2034-2037)

    def _multi_scores(self, bin_indices):
        """[n, outputs] = base + every tree's lr * leaf vector, tree by tree
        (core.py:1904-1935)."""
        hf = self._heap_forest()
        if hf is not None:
            return hf.predict_tree_order(bin_indices)
        n, M = bin_indices.size(0), self.num_outputs
        out = torch.tensor(self._base_vector(), dtype=torch.float32, device=self.device).reshape(1, M).repeat(n, 1)
        bic = bin_indices.contiguous()
        for m, tt in enumerate(self._get_cached_multi_regression_tree_tensors()):
            if tt.size(0) == 0:
                continue
            col = out[:, m].contiguous()
            _nk.predict_forest(bic, tt.contiguous(), self.learning_rate, col)
            out[:, m] = col
        return out

    def _get_cached_multi_regression_tree_tensors(self):
        """One flattened forest per output (core.py:688-731)."""
        max_nodes = 2 ** (self.max_depth + 1)
        if (self._cached_multi_regression_tree_tensors is not None
                and self._prediction_cache_matches("multi_regression", max_nodes, num_outputs=self.num_outputs)):
            return self._cached_multi_regression_tree_tensors
        trees = [t for t in self.forest[: self._trees_trained] if t]
        out = []
        for m in range(self.num_outputs):
            if trees:
                out.append(torch.stack([self.flatten_tree(t, max_nodes, output_index=m) for t in trees]).to(
                    self.device))
            else:
                out.append(torch.empty((0, max_nodes, 6), dtype=torch.float32, device=self.device))
        self._cached_multi_regression_tree_tensors = out
        self._cached_regression_tree_tensor = None
        self._cached_multiclass_tree_tensors = None
        self._set_prediction_cache_meta("multi_regression", max_nodes, num_outputs=self.num_outputs)
        return out

    def _heap_forest(self):
        """The trained forest compiled for the device (cached per tree count):
        from the batched engine's node tables when they cover every trained
        round, else from the nested tree dicts (per-node engine, loaded or
        warm-started models).  None for trees deeper than 16 levels."""
        from .serving import HeapForest
        if self.max_depth > 16 or not self._trees_trained:
            return None
        key = (self._trees_trained, str(self.device))
        hc = getattr(self, "_heap_cache", None)
        if hc is not None and hc[0] == key:
            return hc[1]
        K = self._n_channels()
        base = self._base_vector()
        tabs = getattr(self, "_round_tables", None)
        if tabs and sum(t["status"].shape[0] for t in tabs) == self._trees_trained:
            NH = tabs[0]["status"].shape[-1]
            cat = {k: np.concatenate([t[k] for t in tabs]).reshape(-1, NH) for k in ("status", "feat", "bin", "value")}
            hf = HeapForest.from_tables(cat["status"], cat["feat"], cat["bin"], cat["value"], self.max_depth,
                                        self.learning_rate, self.num_features, base, K, device=self.device)
        else:
            hf = HeapForest.from_model(self)
        self._heap_cache = (key, hf)
        return hf

    def predict_binned(self, bin_indices):
        """base + sum_t lr*leaf in tree order (core.py:1902-1962): the
        compiled heap forest, so no per-tree host flattening; trees deeper
        than the heap forest's 16 levels use the flattened-forest op.
        Multi-output models return [n, outputs]."""
        if self.objective == "multi_regression":
            return self._multi_scores(bin_indices)
        hf = self._heap_forest()
        if hf is not None:
            return hf.predict_tree_order(bin_indices)[:, 0]
        tt = self._get_cached_regression_tree_tensor()
        out = torch.full((bin_indices.size(0),), np.float32(self.base_prediction), device=self.device,
                         dtype=torch.float32)
        if tt.size(0) > 0:
            _nk.predict_forest(bin_indices.contiguous(), tt.contiguous(), self.learning_rate, out)
        return out

    def predict(self, X_np):
        """Regression values or class labels (core.py:1985-2004)."""
        if self.objective in ("multiclass", "binary"):
            probs = self.predict_proba(X_np)
            return self.label_encoder.inverse_transform(np.argmax(probs, axis=1))
        bi = self.bin_inference_data(X_np)
        preds = self.predict_binned(bi).cpu().numpy()
        return preds if self.objective == "multi_regression" else preds.reshape(-1)

    def predict_proba(self, X_np):
        """Class probabilities (core.py:2006-2019)."""
        if self.objective not in ("multiclass", "binary"):
            raise ValueError("predict_proba only available for classification objectives")
        bi = self.bin_inference_data(X_np)
        return self.predict_proba_binned(bi).cpu().numpy()

    def predict_proba_binned(self, bin_indices):
        """Uniform log(1/K) prior + per-class forests, softmax (core.py:2021-2058)."""
        return softmax(self._raw_multiclass_scores(bin_indices), dim=1)

    def _raw_multiclass_scores(self, bin_indices):
        """Per-class raw scores: the log(1/K) prior plus each class's forest."""
        hf = self._heap_forest()
        if hf is not None:
            return hf.predict_tree_order(bin_indices)
        n = bin_indices.size(0)
        K = self.num_classes
        F = torch.zeros((n, K), device=self.device, dtype=torch.float32)
        prior = float(np.log(np.float32(1.0 / K) + 1e-10))
        for k in range(K):
            F[:, k] = prior
        tts = self._get_cached_multiclass_tree_tensors()
        bic = bin_indices.contiguous()
        for k, tt in enumerate(tts):
            if tt is None or tt.size(0) == 0:
                continue
            col = torch.zeros(n, device=self.device, dtype=torch.float32)
            _nk.predict_forest(bic, tt.contiguous(), self.learning_rate, col)
            F[:, k] += col
        return F

    # ------------------------------------------------------------ importance
    def get_feature_importance(self, importance_type="gain", normalize=True):
        """core.py:2060-2093."""
        if self.feature_importance_ is None:
            raise ValueError("Model has not been fitted yet.")
        if importance_type != "gain":
            raise ValueError(f"importance_type '{importance_type}' not supported. Use 'gain'.")
        imp = self.feature_importance_.copy()
        if normalize and imp.sum() > 0:
            imp = imp / imp.sum()
        return imp

    def get_per_era_feature_importance(self, normalize=True):
        """core.py:2095-2129."""
        if self.per_era_feature_importance_ is None:
            raise ValueError("Model has not been fitted yet.")
        imp = self.per_era_feature_importance_.copy()
        if normalize:
            for e in range(imp.shape[0]):
                s = imp[e].sum()
                if s > 0:
                    imp[e] /= s
        return imp

    # ------------------------------------------------------------ save / load
    _STATE_KEYS = ("objective", "num_bins", "max_depth", "learning_rate", "n_estimators", "min_child_weight",
                   "min_split_gain", "L2_reg", "colsample_bytree", "random_state", "warm_start",
                   "monotonic_constraints")

    def save_model(self, path):
        """Pickle the reference's state dict (core.py:2131-2189)."""
        if not self._is_fitted:
            raise ValueError("Cannot save unfitted model. Call fit() first.")
        state = {k: getattr(self, k) for k in self._STATE_KEYS}
        state.update({
            "forest": self.forest, "bin_edges": self.bin_edges,
            "base_prediction": (self.base_prediction.detach().cpu() if torch.is_tensor(self.base_prediction)
                                else self.base_prediction),
            "num_features": self.num_features, "num_classes": self.num_classes, "num_outputs": self.num_outputs,
            "classes_": self.classes_, "label_encoder": self.label_encoder,
            "feature_importance_": self.feature_importance_,
            "per_era_feature_importance_": self.per_era_feature_importance_, "_is_fitted": self._is_fitted,
            "_trees_trained": self._trees_trained, "training_loss": self.training_loss,
            "eval_loss": self.eval_loss, "monotonic_constraints_vector_": self.monotonic_constraints_vector_,
            "monotonic_required_split_direction_": self.monotonic_required_split_direction_,
            "_monotonic_constraints_fitted_": self._monotonic_constraints_fitted_,
        })
        with open(path, "wb") as f:
            pickle.dump(state, f)
        print(f"Model saved to {path}")

    def load_model(self, path):
        """Restore a state dict written by save_model (core.py:2191-2281).

        Model files are pickles: load only files you trust."""
        with open(path, "rb") as f:
            state = pickle.load(f)
        for k in self._STATE_KEYS:
            setattr(self, k, state.get(k, getattr(self, k, None)))
        self.forest = state["forest"]
        self.bin_edges = state["bin_edges"]
        self.base_prediction = state["base_prediction"]
        if self.objective == "multi_regression":  # a per-output vector on the device (core.py:2259-2270)
            self.base_prediction = torch.as_tensor(self.base_prediction, dtype=torch.float32).reshape(-1).to(
                self.device)
        elif torch.is_tensor(self.base_prediction):
            self.base_prediction = float(self.base_prediction.item())
        self.num_features = state["num_features"]
        self.num_classes = state["num_classes"]
        self.num_outputs = state.get("num_outputs", 1 if self.objective == "regression" else None)
        self.classes_ = state["classes_"]
        self.label_encoder = state["label_encoder"]
        self.feature_importance_ = state["feature_importance_"]
        self.per_era_feature_importance_ = state["per_era_feature_importance_"]
        self._is_fitted = state["_is_fitted"]
        self._trees_trained = state["_trees_trained"]
        self.training_loss = state.get("training_loss", [])
        self.eval_loss = state.get("eval_loss", [])
        self.monotonic_constraints_vector_ = state.get("monotonic_constraints_vector_", None)
        self.monotonic_required_split_direction_ = state.get("monotonic_required_split_direction_", None)
        self._monotonic_constraints_fitted_ = state.get("_monotonic_constraints_fitted_",
                                                        self.monotonic_constraints_vector_ is not None)
        self.monotonic_required_split_direction_gpu_ = None
        self._feature_ids = None
        self._round_tables = None  # a loaded forest is compiled from its tree dicts
        self._invalidate_prediction_cache()
        print(f"Model loaded from {path} ({self._trees_trained} trees)")
        return self

    # ------------------------------------------------------------ flatten
    def _feature_of(self, x):
        """Integer feature of a split; a 0-d view into this model's feature-id
        vector is read from its storage offset (no device round trip)."""
        if torch.is_tensor(x):
            ids = self._feature_ids
            if (ids is not None and x.dim() == 0 and x.device == ids.device
                    and x.untyped_storage().data_ptr() == ids.untyped_storage().data_ptr()):
                return int(x.storage_offset())
            return int(x.item())
        return int(x)

    def flatten_tree(self, tree, max_nodes, output_index=None):
        """Preorder [max_nodes, 6] float32 layout (warpgbm/core.py:2283-2328)."""
        rows = []
        counter = [0]

        def walk(node):
            cur = counter[0]
            counter[0] += 1
            if "leaf_value" in node:
                v = node["leaf_value"]
                if isinstance(v, torch.Tensor):
                    v = v.detach().cpu().tolist()
                if isinstance(v, (list, tuple, np.ndarray)):
                    v = v[0 if output_index is None else output_index]
                rows.append((cur, (np.nan, np.nan, np.nan, np.nan, 1.0, float(v))))
            else:
                f = float(self._feature_of(node["feature"]))
                b = float(node["bin"])
                lid = counter[0]
                walk(node["left"])
                rid = counter[0]
                walk(node["right"])
                rows.append((cur, (f, b, float(lid), float(rid), 0.0, np.nan)))

        walk(tree)
        flat = np.full((max_nodes, 6), np.nan, dtype=np.float32)
        for i, r in rows:
            flat[i] = r
        return torch.from_numpy(flat)
