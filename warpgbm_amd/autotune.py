"""Autotuner for the histogram kernel's tiling.

For a histogram shape (features, bins, matrix columns) it times candidate
configurations -- rows per chunk C, waves per workgroup W (the waves of a
workgroup share one LDS histogram), features per lane FPL (load width and
LDS footprint) and load depth L -- on synthetic codes of that shape with
the engine's own histogram op, and installs the fastest with
``wgbm_set_hist_config``.  Results are cached per shape in memory and in a
JSON file (``$WARPGBM_AMD_TUNE_CACHE``, default
``~/.cache/warpgbm_amd/hist_tune.json``).

Histograms are exact fixed-point sums, so every configuration yields the
same bits: tuning changes speed only, never a tree.
"""
from __future__ import annotations

import ctypes
import itertools
import json
import os
from pathlib import Path

import torch

from . import _lib

_MEM: dict = {}


def _cache_path() -> Path:
    p = os.environ.get("WARPGBM_AMD_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "warpgbm_amd" / "hist_tune.json"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except Exception:
        return {}


def _save_cache(d: dict) -> None:
    try:
        p = _cache_path()
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(d, indent=1, sort_keys=True))
    except OSError:
        pass


def lds_bytes(fpl: int, n_bins: int, planes: int) -> int:
    """Shared histogram of one workgroup: planes x [fpl][bins + 1][64] u64."""
    return planes * fpl * (n_bins + 1) * 64 * 8


def candidates(n_feat: int, n_bins: int, n_cols: int, const_hessian: bool = False):
    """Configurations whose shared histogram fits a workgroup's 160 KB LDS
    (two planes unless every row has the same hessian)."""
    out = []
    planes = 1 if const_hessian else 2
    for fpl, depths in ((1, (32, 64)), (2, (16, 32)), (4, (8, 16)), (8, (4, 8)), (16, (2, 4))):
        if n_cols % fpl or lds_bytes(fpl, n_bins, planes) > 160 * 1024:
            continue
        if fpl > 1 and 64 * fpl > 2 * max(n_feat, 1):
            continue  # slab far wider than the features
        for c, w, d in itertools.product((4096, 8192, 16384), (8, 16) if fpl <= 4 else (4, 8), depths):
            out.append((c, w, fpl, d))
    return out


def current(n_feat: int, n_bins: int, n_cols: int, const_hessian: bool = False):
    out = (ctypes.c_int32 * 4)()
    _lib.call("wgbm_get_hist_config", n_feat, n_bins, n_cols, int(const_hessian), out)
    return tuple(out)


def apply(cfg) -> None:
    c, w, f, d = cfg
    _lib.call("wgbm_set_hist_config", int(c), int(w), int(f), int(d))


def reset() -> None:
    _lib.call("wgbm_set_hist_config", 0, 0, 0, 0)


def tune(n_feat: int, n_bins: int, n_cols: int | None = None, rows: int = 262_144, device="cuda",
         repeats: int = 2, use_cache: bool = True, verbose: bool = False):
    """Pick the fastest histogram configuration for this shape and install it.

    Returns (config, {config: ms}).  The timing builds one histogram of
    ``rows`` rows over all ``n_feat`` features (the root build of a tree).
    """
    from .cuda import node_kernel as nk
    n_cols = n_feat if n_cols is None else n_cols
    key = f"{n_feat}x{n_bins}x{n_cols}@{torch.cuda.get_device_name(device)}"
    disk = _load_cache() if use_cache else {}
    if use_cache and key in _MEM:
        apply(_MEM[key])
        return _MEM[key], {}
    if use_cache and key in disk:
        cfg = tuple(disk[key]["config"])
        _MEM[key] = cfg
        apply(cfg)
        return cfg, {}
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    codes = torch.randint(0, n_bins, (rows, n_cols), generator=g, device=dev, dtype=torch.int8)
    gr = torch.randn(rows, generator=g, device=dev)
    hs = torch.rand(rows, generator=g, device=dev) + 0.5
    idx = torch.arange(rows, dtype=torch.int32, device=dev)
    feats = torch.arange(n_feat, dtype=torch.int32, device=dev)
    eras = torch.zeros(rows, dtype=torch.int32, device=dev)
    G = torch.zeros((1, n_feat, n_bins), device=dev)
    H = torch.zeros_like(G)
    times = {}
    for cfg in candidates(n_feat, n_bins, n_cols):
        apply(cfg)
        nk.compute_histogram3(codes, gr, hs, idx, feats, eras, G, H, n_bins)  # warm-up
        torch.cuda.synchronize(dev)
        _lib.call("wgbm_profile_begin")
        for _ in range(repeats):
            nk.compute_histogram3(codes, gr, hs, idx, feats, eras, G, H, n_bins)
        out = (ctypes.c_double * 4)()
        _lib.call("wgbm_profile_end", out)
        times[cfg] = (out[0] + out[2]) / repeats
        if verbose:
            print(cfg, f"{times[cfg]:.3f} ms")
    best = min(times, key=times.get)
    apply(best)
    _MEM[key] = best
    if use_cache:
        disk[key] = {"config": list(best), "ms": times[best]}
        _save_cache(disk)
    return best, times
