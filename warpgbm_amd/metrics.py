"""Evaluation metrics and softmax over PyTorch tensors.

The softmax and the weighted metrics the estimator's evaluation uses, with
the formulas of warpgbm/metrics.py (softmax :12-15, weighted metrics
:46-50, :86-127).  These run once per evaluation round on the host's
orchestration side; they are not part of the training hot path.
"""
from __future__ import annotations

import torch


def softmax(logits, dim=-1):
    """Max-subtracted softmax (warpgbm/metrics.py:12-15)."""
    shifted = logits - torch.max(logits, dim=dim, keepdim=True)[0]
    e = torch.exp(shifted)
    return e / torch.sum(e, dim=dim, keepdim=True)


def _weights(w, like):
    return torch.ones_like(like, dtype=torch.float32) if w is None else w.to(dtype=torch.float32)


def weighted_mse_torch(y_true, y_pred, sample_weight=None):
    w = _weights(sample_weight, y_true)
    return torch.sum(w * (y_true - y_pred) ** 2) / torch.sum(w)


def weighted_rmsle_torch(y_true, y_pred, sample_weight=None, eps=1e-7):
    w = _weights(sample_weight, y_true)
    lt = torch.log1p(torch.clamp(y_true, min=0) + eps)
    lp = torch.log1p(torch.clamp(y_pred, min=0) + eps)
    return torch.sqrt(torch.sum(w * (lt - lp) ** 2) / torch.sum(w))


def weighted_log_loss_torch(y_true_labels, y_pred_probs, sample_weight=None, eps=1e-15):
    p = torch.clamp(y_pred_probs, eps, 1 - eps)
    idx = torch.arange(y_true_labels.shape[0], device=y_true_labels.device)
    w = _weights(sample_weight, y_true_labels)
    return torch.sum(w * -torch.log(p[idx, y_true_labels.long()])) / torch.sum(w)


def weighted_accuracy_torch(y_true_labels, y_pred_labels, sample_weight=None):
    w = _weights(sample_weight, y_true_labels)
    return torch.sum(w * (y_true_labels == y_pred_labels).float()) / torch.sum(w)


def weighted_corr_loss_torch(y_true, y_pred, sample_weight=None, eps=1e-12):
    w = _weights(sample_weight, y_true)
    ws = torch.sum(w)
    ct = y_true - torch.sum(w * y_true) / ws
    cp = y_pred - torch.sum(w * y_pred) / ws
    cov = torch.sum(w * ct * cp) / ws
    vt = torch.sum(w * ct * ct) / ws
    vp = torch.sum(w * cp * cp) / ws
    return 1.0 - cov / torch.sqrt(torch.clamp(vt * vp, min=eps))


def weighted_multi_mse_torch(y_true, y_pred, sample_weight=None, valid_mask=None):
    """Weighted MSE over [n, outputs] targets (warpgbm/metrics.py:52-83):
    entries outside valid_mask (default: the finite targets) count with
    weight 0; 1-D weights apply to every output of a row."""
    if y_true.ndim != 2 or y_true.shape != y_pred.shape:
        raise ValueError(f"expected matching [n, outputs] tensors, got {tuple(y_true.shape)} and "
                         f"{tuple(y_pred.shape)}")
    valid = torch.isfinite(y_true) if valid_mask is None else valid_mask.to(torch.bool)
    w = _weights(sample_weight, y_true)
    if w.ndim == 1:
        w = w.view(-1, 1)
    err2 = torch.where(valid, (y_true - y_pred) ** 2, torch.zeros_like(y_pred))
    ew = w * valid.to(torch.float32)
    den = torch.sum(ew)
    if den <= 0:
        raise ValueError("No valid weighted targets available for multi-output MSE.")
    return torch.sum(ew * err2) / den
