"""ctypes binding of the gfx950 HIP library (``libwarpgbm_amd.so``).

The library exports the C ABI declared in ``include/warpgbm_amd.h``.  This
module loads it, declares every signature, and turns non-zero return codes
into ``RuntimeError`` with the library's message.  There is no fallback:
when the library is missing every operation raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libwarpgbm_amd.so"

c_i8p = C.POINTER(C.c_int8)
vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f32 = C.c_float
f64 = C.c_double
sz = C.c_size_t


class FitConfig(C.Structure):
    _fields_ = [
        ("n_rows", i64), ("n_cols", i64), ("n_feat", i32), ("n_bins", i32), ("n_eras", i32),
        ("n_classes", i32), ("max_depth", i32), ("objective", i32), ("learning_rate", f32),
        ("min_split_gain", f32), ("min_child_weight", f32), ("lambda_", f32), ("const_weight", i32),
    ]


class FitIO(C.Structure):
    _fields_ = [
        ("codes", vp), ("y", vp), ("labels", vp), ("weight", vp), ("root_order", vp), ("era_count", vp),
        ("mono_required", vp), ("feat_idx", vp), ("scores", vp), ("node_status", vp), ("node_feat", vp),
        ("node_bin", vp), ("node_value", vp), ("node_count", vp), ("node_era_gain", vp), ("train_loss", vp),
    ]


ABI_VERSION = 4  # wgbm_abi_version() of the library this binding matches

# name -> (restype, argtypes)
SIGNATURES = {
    "wgbm_last_error": (C.c_char_p, []),
    "wgbm_abi_version": (C.c_int, []),
    "wgbm_build_arch": (C.c_char_p, []),
    "wgbm_check_device": (C.c_int, [C.c_char_p, i32]),
    "wgbm_histogram3_workspace_bytes": (sz, [i64, i32, i32, i32, i64]),
    "wgbm_compute_histogram3": (C.c_int, [vp, i64, i64, vp, vp, vp, i64, vp, i32, vp, vp, vp, i32, i32, vp, sz, vp]),
    "wgbm_compute_split": (C.c_int, [vp, vp, i32, i32, i32, f32, f32, f32, vp, vp, vp]),
    "wgbm_bin_column": (C.c_int, [vp, i64, vp, i32, vp, vp]),
    "wgbm_predict_forest": (C.c_int, [vp, i64, i64, vp, i64, i64, f32, vp, vp]),
    "wgbm_select_split": (C.c_int, [vp, vp, i32, i32, i32, f32, vp, vp, vp, vp]),
    "wgbm_partition_workspace_bytes": (sz, [i64]),
    "wgbm_partition_rows": (C.c_int, [vp, i64, i64, vp, i64, i32, i32, vp, vp, vp, sz, vp]),
    "wgbm_leaf_workspace_bytes": (sz, [i64]),
    "wgbm_leaf_value": (C.c_int, [vp, vp, vp, i64, i64, f64, f32, vp, i32, i32, vp, vp, vp, sz, vp]),
    "wgbm_row_stats_workspace_bytes": (sz, []),
    "wgbm_row_stats": (C.c_int, [vp, vp, vp, i64, i64, vp, vp, sz, vp]),
    "wgbm_hist_subtract": (C.c_int, [vp, vp, vp, vp, vp, vp, i64, vp]),
    "wgbm_grad_regression": (C.c_int, [vp, vp, vp, i64, vp, vp, vp]),
    "wgbm_grad_softmax": (C.c_int, [vp, vp, vp, i64, i32, vp, vp, vp]),
    "wgbm_grad_multi_regression": (C.c_int, [vp, vp, vp, i64, i32, vp, vp, vp, vp, vp]),
    "wgbm_edges_from_sorted": (C.c_int, [vp, i64, i32, i32, vp, i32, vp]),
    "wgbm_bin_matrix_workspace_bytes": (sz, [i32, i32]),
    "wgbm_bin_matrix": (C.c_int, [vp, i64, i32, vp, i32, vp, vp, sz, vp]),
    "wgbm_set_hist_config": (C.c_int, [i32, i32, i32, i32, i32]),
    "wgbm_set_hist_config_shape": (C.c_int, [i32, i32, i64, i32, i32, i32, i32, i32, i32]),
    "wgbm_get_hist_config": (C.c_int, [i32, i32, i64, i32, C.POINTER(i32)]),
    "wgbm_profile_begin": (C.c_int, []),
    "wgbm_profile_end": (C.c_int, [C.POINTER(f64)]),
    "wgbm_fit_workspace_bytes": (sz, [C.POINTER(FitConfig)]),
    "wgbm_fit_rounds": (C.c_int, [C.POINTER(FitConfig), C.POINTER(FitIO), i32, vp, sz, vp]),
    "wgbm_fit_graph_rounds": (i32, []),
    "wgbm_set_fit_graph": (C.c_int, [i32]),
    "wgbm_heap_compile": (C.c_int, [vp, vp, vp, vp, i32, i32, i32, f32, vp, vp, vp]),
    "wgbm_heap_predict_workspace_bytes": (sz, [i64, i32, i32]),
    "wgbm_set_predict_mode": (C.c_int, [i32]),
    "wgbm_quantile_edges_workspace_bytes": (sz, [i64, i32]),
    "wgbm_quantile_edges": (C.c_int, [vp, i64, i32, i64, i32, vp, i32, vp, sz, vp]),
    "wgbm_heap_predict": (C.c_int, [vp, i64, i32, vp, vp, i32, i32, i32, vp, vp, i32, i32, vp, vp, vp, sz, vp]),
    "wgbm_heap_predict_tree_order": (C.c_int, [vp, i64, i32, vp, vp, i32, i32, vp, i32, vp, vp, vp]),
    "wgbm_heap_predict_noalloc": (C.c_int, [vp, i64, i32, vp, vp, i32, i32, i32, vp, vp, i32, i32, vp, vp, vp]),
}

_lib = None
_load_error = None


def lib_path() -> Path:
    return _LIB_PATH


def load():
    """Load the library once; raise ImportError if it is missing or stale."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise ImportError(_load_error)
    if not _LIB_PATH.exists():
        _load_error = (f"warpgbm_amd HIP library not found at {_LIB_PATH}; build it with "
                       "`python -m warpgbm_amd.build` (hipcc --offload-arch=gfx950)")
        raise ImportError(_load_error)
    try:
        lib = C.CDLL(str(_LIB_PATH))
    except OSError as e:  # pragma: no cover - depends on the system
        _load_error = f"cannot load {_LIB_PATH}: {e}"
        raise ImportError(_load_error) from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.wgbm_abi_version() != ABI_VERSION:
        _load_error = (f"{_LIB_PATH} has ABI {lib.wgbm_abi_version()}, this package needs {ABI_VERSION}; "
                       "rebuild with `python -m warpgbm_amd.build`")
        raise ImportError(_load_error)
    _lib = lib
    return lib


def call(name: str, *args):
    """Call an int-returning entry point and raise on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.wgbm_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))
