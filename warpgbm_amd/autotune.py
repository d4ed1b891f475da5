"""Autotuner for the histogram kernel's tiling.

A configuration is rows per chunk C, waves per workgroup W (the waves of a
workgroup share one LDS histogram), features per lane FPL (load width and
LDS footprint, i.e. how many sub-histograms a workgroup holds), load depth
L (rows whose code loads are in flight per wave) and slots S (workgroups a
launch aims at: fewer slots mean fewer chunk partials to write and merge).  These take over
from the reference's threads_per_block / rows_per_thread launch knobs
(warpgbm/core.py:68-69, :200-207).

* ``tune_fit`` (called by ``WarpGBM.fit`` on the batched engine): times
  every candidate on a few boosting rounds of the fit's own data -- the
  level builds of every depth, with the variant that fit runs (one LDS plane
  when weights are equal) -- and installs the fastest for that histogram
  shape only (``wgbm_set_hist_config_shape``).
* ``tune``: times candidates on one synthetic root build with the per-node
  histogram op and installs the winner process-wide (``wgbm_set_hist_config``).

Results are cached per shape and device in memory and in a JSON file
(``$WARPGBM_AMD_TUNE_CACHE``, default ``~/.cache/warpgbm_amd/hist_tune.json``).
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
    """Configurations whose shared histogram fits a workgroup's LDS (160 KB
    less the kernel's own 1 KB; two planes unless every row has the same
    hessian)."""
    out = []
    planes = 1 if const_hessian else 2
    for fpl, depths in ((1, (32, 64)), (2, (16, 32)), (4, (8, 16)), (8, (4, 8)), (16, (2, 4))):
        if n_cols % fpl or lds_bytes(fpl, n_bins, planes) > 159 * 1024:
            continue
        if fpl > 1 and 64 * fpl > 2 * max(n_feat, 1):
            continue  # slab far wider than the features
        for c, w, d, s in itertools.product((4096, 8192, 16384), (8, 16) if fpl <= 4 else (4, 8), depths,
                                            (256, 512)):
            out.append((c, w, fpl, d, s))
    return out


def current(n_feat: int, n_bins: int, n_cols: int, const_hessian: bool = False):
    out = (ctypes.c_int32 * 5)()
    _lib.call("wgbm_get_hist_config", n_feat, n_bins, n_cols, int(const_hessian), out)
    return tuple(out)


def apply(cfg) -> None:
    c, w, f, d, s = cfg
    _lib.call("wgbm_set_hist_config", int(c), int(w), int(f), int(d), int(s))


def reset() -> None:
    _lib.call("wgbm_set_hist_config", 0, 0, 0, 0, 0)


def tune(n_feat: int, n_bins: int, n_cols: int | None = None, rows: int = 262_144, device="cuda",
         repeats: int = 2, use_cache: bool = True, verbose: bool = False):
    """Pick the fastest histogram configuration for this shape and install it.

    Returns (config, {config: ms}).  The timing builds one histogram of
    ``rows`` rows over all ``n_feat`` features (the root build of a tree).
    """
    from .cuda import node_kernel as nk
    n_cols = n_feat if n_cols is None else n_cols
    key = f"op2:{n_feat}x{n_bins}x{n_cols}@{torch.cuda.get_device_name(device)}"
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


def tune_fit(model, cfg, n_feat, rounds: int = 2, repeats: int = 3, use_cache: bool = True, verbose=False,
             warm_rounds: int = 8):
    """Pick the histogram configuration of a batched fit and install it for
    the fit's histogram shape.  ``model``: a WarpGBM whose training data are
    set up on the device; ``cfg``: its wgbm_fit_config.  The scores are first
    advanced ``warm_rounds`` rounds on a copy (fits of at least four times
    that many trees), so that candidates are timed on trees like the fit's
    typical ones rather than its first; each candidate then runs ``rounds``
    rounds from there in one call (``WarpGBM._trial_rounds_ms``; fewer than
    three rounds launch eagerly, no graph to capture), timed on the device
    as a whole: every launch of the call, not only the histogram ones, since
    a tiling also moves the cost of the merges.  The best of ``repeats``
    runs counts.  The model's scores are left as they were.  Returns a dict
    describing the choice."""
    import time
    one_plane = cfg.objective == 0 and cfg.const_weight != 0
    B, n_cols = int(cfg.n_bins), int(cfg.n_cols)
    dev = torch.device(model.device)
    key = f"fit3:{n_feat}x{B}x{n_cols}:{'one' if one_plane else 'two'}-plane@{torch.cuda.get_device_name(dev)}"
    hit = _MEM.get(key) if use_cache else None
    disk = _load_cache() if use_cache and hit is None else {}
    if hit is None and use_cache and key in disk:
        hit = dict(disk[key], config=tuple(disk[key]["config"]))
    if hit is not None:
        _lib.call("wgbm_set_hist_config_shape", n_feat, B, n_cols, int(one_plane), *hit["config"])
        _MEM[key] = hit
        return dict(hit, cached=True)
    t0 = time.perf_counter()
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    _lib.call("wgbm_set_hist_config_shape", n_feat, B, n_cols, int(one_plane), 0, 0, 0, 0, 0)
    auto = current(n_feat, B, n_cols, one_plane)
    saved = (model.gradients, getattr(model, "_last_tables", None))
    times = {}
    try:
        if warm_rounds > 0 and model.n_estimators >= 4 * warm_rounds:
            model.gradients = saved[0].clone()
            model._advance_rounds(cfg, warm_rounds, gen)
        model._trial_rounds_ms(cfg, 1, gen)  # warm-up: first launches of the kernels
        for c in [auto] + candidates(n_feat, B, n_cols, one_plane):
            _lib.call("wgbm_set_hist_config_shape", n_feat, B, n_cols, int(one_plane), *c)
            c = current(n_feat, B, n_cols, one_plane)  # the configuration in effect (depths are clamped)
            if c in times:
                continue
            times[c] = min(model._trial_rounds_ms(cfg, rounds, gen) for _ in range(repeats)) / rounds
            if verbose:
                print(c, f"{times[c]:.3f} ms per round")
    finally:
        model.gradients, model._last_tables = saved
    best = min(times, key=times.get)
    _lib.call("wgbm_set_hist_config_shape", n_feat, B, n_cols, int(one_plane), *best)
    res = {"config": tuple(best), "ms_per_round": times[best], "auto_config": tuple(auto),
           "auto_ms_per_round": times[auto], "candidates": len(times), "tune_s": time.perf_counter() - t0}
    _MEM[key] = res
    if use_cache:
        disk[key] = dict(res, config=list(best), auto_config=list(auto))
        _save_cache(disk)
    return dict(res, cached=False)
