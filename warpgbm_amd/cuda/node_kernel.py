"""The reference's operator module, served by the gfx950 HIP library.

Drop-in replacement for ``warpgbm.cuda.node_kernel`` (the pybind11 module of
warpgbm/cuda/node_kernel.cpp:36-41): the same four functions with the same
positional arguments and in-place semantics, over PyTorch device tensors.
Each call validates dtypes/shapes (raising ``RuntimeError`` like a failed
``TORCH_CHECK``), then calls the C ABI on the current PyTorch HIP stream.
Unlike the reference, launch errors raise instead of printing.
"""
from __future__ import annotations

import torch

from .. import _lib


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _check(t: torch.Tensor, dtype: torch.dtype, name: str, ndim: int | None = None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a torch.Tensor")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be a device tensor, got device {t.device}")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if ndim is not None and t.dim() != ndim:
        raise RuntimeError(f"{name} must be {ndim}-dimensional, got shape {tuple(t.shape)}")


# Workspaces kept between calls, one per (device, stream): a call reuses the
# buffer of the previous call on its stream (which the stream has finished
# with by the time this call's kernels run) and only grows it.  The largest
# workspace of each stream is held until the process ends.
_WORKSPACES: dict = {}


def _workspace(device, nbytes: int, stream: int) -> torch.Tensor:
    key = (str(device), stream)
    ws = _WORKSPACES.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
        _WORKSPACES[key] = ws
    return ws


def compute_histogram3(bin_indices, grads, hess, sample_indices, feature_indices, era_indices,
                       grad_hist, hess_hist, num_bins, threads_per_block=256, rows_per_thread=1):
    """Per-(era, feature, bin) sums of grads/hess over the listed rows.

    Same contract as the reference (warpgbm/cuda/histogram_kernel.cu:64-107):
    histograms [E, F, B] float32 accumulate; E is read from grad_hist.size(0);
    codes outside [0, num_bins) are dropped.  threads_per_block and
    rows_per_thread are accepted for signature compatibility; tiling comes
    from the engine's histogram configuration.
    """
    _check(bin_indices, torch.int8, "bin_indices", 2)
    _check(grads, torch.float32, "grads")
    _check(hess, torch.float32, "hess")
    _check(sample_indices, torch.int32, "sample_indices", 1)
    _check(feature_indices, torch.int32, "feature_indices", 1)
    _check(era_indices, torch.int32, "era_indices")
    _check(grad_hist, torch.float32, "grad_hist", 3)
    _check(hess_hist, torch.float32, "hess_hist", 3)
    n_eras, n_feat, nb = grad_hist.shape
    if tuple(hess_hist.shape) != tuple(grad_hist.shape):
        raise RuntimeError("grad_hist and hess_hist shapes differ")
    if n_feat != feature_indices.numel() or nb != int(num_bins):
        raise RuntimeError(f"histogram shape {tuple(grad_hist.shape)} does not match "
                           f"[E, {feature_indices.numel()}, {int(num_bins)}]")
    n_rows, n_cols = bin_indices.shape
    if grads.numel() < n_rows or hess.numel() < n_rows or era_indices.numel() < n_rows:
        raise RuntimeError("grads, hess and era_indices need one value per row of bin_indices")
    n = sample_indices.numel()
    st = _stream(bin_indices)
    ws = _workspace(bin_indices.device,
                    _lib.query("wgbm_histogram3_workspace_bytes", n, n_feat, int(num_bins), n_eras, n_cols), st)
    _lib.call("wgbm_compute_histogram3", bin_indices.data_ptr(), n_rows, n_cols, grads.data_ptr(),
              hess.data_ptr(), sample_indices.data_ptr(), n, feature_indices.data_ptr(), n_feat,
              era_indices.data_ptr(), grad_hist.data_ptr(), hess_hist.data_ptr(), n_eras, int(num_bins),
              ws.data_ptr(), ws.numel(), st)


def compute_split(G, H, min_split_gain, min_child_samples, eps, per_era_gain, per_era_direction, threads=128):
    """Per-era gain and direction, warpgbm/cuda/best_split_kernel.cu:57-88."""
    _check(G, torch.float32, "G", 3)
    _check(H, torch.float32, "H", 3)
    _check(per_era_gain, torch.float32, "per_era_gain", 3)
    _check(per_era_direction, torch.float32, "per_era_direction", 3)
    E, F, B = G.shape
    if tuple(H.shape) != (E, F, B):
        raise RuntimeError("G and H shapes differ")
    if tuple(per_era_gain.shape) != (E, F, B - 1) or tuple(per_era_direction.shape) != (E, F, B - 1):
        raise RuntimeError(f"gain/direction must be [{E}, {F}, {B - 1}]")
    _lib.call("wgbm_compute_split", G.data_ptr(), H.data_ptr(), E, F, B, float(min_split_gain),
              float(min_child_samples), float(eps), per_era_gain.data_ptr(), per_era_direction.data_ptr(),
              _stream(G))


def custom_cuda_binner(X, bin_edges, bin_indices):
    """int8 codes of one float32 column, warpgbm/cuda/binner.cu:29-52."""
    _check(X, torch.float32, "X", 1)
    _check(bin_edges, torch.float32, "bin_edges", 1)
    _check(bin_indices, torch.int8, "bin_indices", 1)
    if bin_indices.numel() != X.numel():
        raise RuntimeError("bin_indices must have one entry per value of X")
    _lib.call("wgbm_bin_column", X.data_ptr(), X.numel(), bin_edges.data_ptr(), bin_edges.numel(),
              bin_indices.data_ptr(), _stream(X))


def predict_forest(bin_indices, tree_tensor, learning_rate, out):
    """out += lr * leaf over a flattened forest, warpgbm/cuda/predict.cu:48-76."""
    _check(bin_indices, torch.int8, "bin_indices", 2)
    _check(tree_tensor, torch.float32, "tree_tensor", 3)
    _check(out, torch.float32, "out", 1)
    N, F = bin_indices.shape
    T, M, six = tree_tensor.shape
    if six != 6:
        raise RuntimeError("tree_tensor must be [T, max_nodes, 6]")
    if out.numel() != N:
        raise RuntimeError("out must have one entry per row")
    _lib.call("wgbm_predict_forest", bin_indices.data_ptr(), N, F, tree_tensor.data_ptr(), T, M,
              float(learning_rate), out.data_ptr(), _stream(out))
