"""ctypes wrapper of the multithreaded FP64 C restatement (cpu_gbm.c).

TEST INFRASTRUCTURE ONLY (the timed CPU baseline of bench.py and a checker
in tests/).  ``load()`` builds ``libgbm_cpu.so`` with ``make -C oracle``
when it is missing.  Trees come back as heap tables (root 0, children
2i+1 / 2i+2; status 0 absent, 1 split, 2 leaf) and as nested dicts in the
layout of gbm_oracle.grow_tree.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_DIR, "libgbm_cpu.so")
_lib = None

_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double


def build(force=True):
    """Compile libgbm_cpu.so (force: even when it is up to date)."""
    subprocess.run(["make", "-s", "-C", _DIR] + (["-B"] if force else []), check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(force=False)
        lib = C.CDLL(_LIB)
        lib.gbm_cpu_grow_tree.restype = C.c_int
        lib.gbm_cpu_grow_tree.argtypes = [_P, _I64, _I32, _P, _P, _P, _I32, _I32, _I32, _D, _D, _D, _P, _I32,
                                          _P, _P, _P, _P, _P, _P]
        lib.gbm_cpu_fit_regression.restype = C.c_int
        lib.gbm_cpu_fit_regression.argtypes = [_P, _I64, _I32, _P, _P, _P, _I32, _I32, _I32, _I32, _D, _D, _D, _D,
                                               _P, _I32, _P, _P, _P, _P, _P, _P, _P]
        lib.gbm_cpu_fit_softmax.restype = C.c_int
        lib.gbm_cpu_fit_softmax.argtypes = [_P, _I64, _I32, _P, _I32, _P, _I32, _I32, _I32, _D, _D, _D, _D, _I32,
                                            _P, _P, _P, _P, _P, _P]
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _tables(shape):
    return (np.zeros(shape, np.int32), np.full(shape, -1, np.int32), np.full(shape, -1, np.int32),
            np.zeros(shape, np.float64), np.zeros(shape, np.int32))


def to_dict(status, feat, binv, value, count, heap=0):
    """Nested dict of one heap table (gbm_oracle.grow_tree's layout)."""
    if status[heap] == 2:
        return {"leaf_value": float(value[heap]), "samples": int(count[heap])}
    return {"feature": int(feat[heap]), "bin": int(binv[heap]),
            "left": to_dict(status, feat, binv, value, count, 2 * heap + 1),
            "right": to_dict(status, feat, binv, value, count, 2 * heap + 2)}


def grow_tree(codes, g, h, eras, num_eras, num_bins, max_depth, min_child_weight, min_split_gain, lam,
              mono_required=None, threads=0):
    """One tree; returns (tables, leaf_of_row).  Same arguments as
    gbm_oracle.grow_tree (all features)."""
    codes = np.ascontiguousarray(codes, np.int8)
    n, F = codes.shape
    g = np.ascontiguousarray(g, np.float64)
    h = np.ascontiguousarray(h, np.float64)
    era = None if num_eras == 1 else np.ascontiguousarray(eras, np.int32)
    mono = None if mono_required is None else np.ascontiguousarray(mono_required, np.int8)
    NH = (2 << max_depth) - 1
    st, ft, bt, vt, ct = _tables(NH)
    leaf = np.zeros(n, np.int32)
    rc = load().gbm_cpu_grow_tree(_ptr(codes), n, F, _ptr(g), _ptr(h), _ptr(era), num_eras, num_bins, max_depth,
                                  float(min_child_weight), float(min_split_gain), float(lam), _ptr(mono),
                                  int(threads), _ptr(st), _ptr(ft), _ptr(bt), _ptr(vt), _ptr(ct), _ptr(leaf))
    if rc != 0:
        raise RuntimeError("gbm_cpu_grow_tree failed")
    return (st, ft, bt, vt, ct), leaf


def fit_regression(codes, y, w, eras, num_eras, num_bins, max_depth, n_trees, lr, min_child_weight=20,
                   min_split_gain=0.0, lam=1e-6, mono_required=None, threads=0):
    """Regression boosting; returns (base, tables [T][NH] each, scores)."""
    codes = np.ascontiguousarray(codes, np.int8)
    n, F = codes.shape
    y = np.ascontiguousarray(y, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    era = None if num_eras == 1 else np.ascontiguousarray(eras, np.int32)
    mono = None if mono_required is None else np.ascontiguousarray(mono_required, np.int8)
    NH = (2 << max_depth) - 1
    st, ft, bt, vt, ct = _tables((n_trees, NH))
    scores = np.zeros(n, np.float64)
    base = C.c_double(0.0)
    rc = load().gbm_cpu_fit_regression(_ptr(codes), n, F, _ptr(y), _ptr(w), _ptr(era), num_eras, num_bins,
                                       max_depth, n_trees, float(lr), float(min_child_weight),
                                       float(min_split_gain), float(lam), _ptr(mono), int(threads),
                                       C.addressof(base), _ptr(st), _ptr(ft), _ptr(bt), _ptr(vt), _ptr(ct),
                                       _ptr(scores))
    if rc != 0:
        raise RuntimeError("gbm_cpu_fit_regression failed")
    return base.value, (st, ft, bt, vt, ct), scores


def fit_softmax(codes, labels, w, num_classes, num_bins, max_depth, n_rounds, lr, min_child_weight=20,
                min_split_gain=0.0, lam=1e-6, threads=0):
    """Softmax boosting rounds; returns (tables [R][K][NH] each, scores [n, K])."""
    codes = np.ascontiguousarray(codes, np.int8)
    n, F = codes.shape
    labels = np.ascontiguousarray(labels, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    NH = (2 << max_depth) - 1
    st, ft, bt, vt, ct = _tables((n_rounds, num_classes, NH))
    scores = np.zeros((n, num_classes), np.float64)
    rc = load().gbm_cpu_fit_softmax(_ptr(codes), n, F, _ptr(labels), num_classes, _ptr(w), num_bins, max_depth,
                                    n_rounds, float(lr), float(min_child_weight), float(min_split_gain), float(lam),
                                    int(threads), _ptr(st), _ptr(ft), _ptr(bt), _ptr(vt), _ptr(ct), _ptr(scores))
    if rc != 0:
        raise RuntimeError("gbm_cpu_fit_softmax failed")
    return (st, ft, bt, vt, ct), scores
