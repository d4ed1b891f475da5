"""Resident heap forests for serving predictions.

A trained forest is compiled once into the device layout the serving
kernel walks (``wgbm_heap_compile``): every tree padded to a complete binary
tree of depth D, nodes packed as (feature << 8 | bin), leaves pre-multiplied
by the learning rate.  ``HeapForest.predict`` then serves any batch of bin
code rows with one traversal launch plus a fixed-order reduction, so a
row's score does not depend on which rows share its launch.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


class HeapForest:
    """Device-resident compiled forest.

    Build with ``from_tables`` (level-order node tables as written by the
    batched trainer) or ``from_model`` (any fitted WarpGBM, via its nested
    tree dicts).
    """

    def __init__(self, nodes, leaves, n_trees, depth, n_feat, class_first, base, tree_block=64, device="cuda"):
        self.nodes = nodes
        self.leaves = leaves
        self.n_trees = int(n_trees)
        self.depth = int(depth)
        self.n_feat = int(n_feat)
        self.device = torch.device(device)
        cf = np.asarray(class_first, dtype=np.int64)
        self.n_classes = len(cf) - 1
        self.tree_block = int(tree_block)
        per_class = np.diff(cf)
        blocks = np.maximum(1, -(-per_class // self.tree_block)).astype(np.int32)
        self.n_blocks = int(blocks.max()) if blocks.size else 1
        self.class_first = torch.from_numpy(cf.astype(np.int32)).to(self.device)
        self.blocks_of_class = torch.from_numpy(blocks).to(self.device)
        self.base = torch.as_tensor(np.asarray(base, dtype=np.float32).reshape(-1), device=self.device)
        self._ws = None

    # -------------------------------------------------------------- builders
    @classmethod
    def from_tables(cls, status, feat, binv, value, depth, learning_rate, n_feat, base, n_classes=1,
                    tree_block=64, device="cuda"):
        """status/feat/bin/value: [T, heap slots] (trees of class k at
        t % n_classes == k, rounds in order)."""
        dev = torch.device(device)
        status = np.ascontiguousarray(status, np.int32)
        T, src = status.shape
        order = np.concatenate([np.arange(k, T, n_classes) for k in range(n_classes)])
        counts = [len(np.arange(k, T, n_classes)) for k in range(n_classes)]
        cf = np.concatenate([[0], np.cumsum(counts)])
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(np.asarray(a)[order], dt)).to(dev)
        st, ft, bt, vt = t(status, np.int32), t(feat, np.int32), t(binv, np.int32), t(value, np.float32)
        internal = (1 << depth) - 1
        nodes = torch.empty((T, internal), dtype=torch.int32, device=dev)
        leaves = torch.empty((T, internal + 1), dtype=torch.float32, device=dev)
        _lib.call("wgbm_heap_compile", st.data_ptr(), ft.data_ptr(), bt.data_ptr(), vt.data_ptr(), src, T, depth,
                  float(learning_rate), nodes.data_ptr(), leaves.data_ptr(), _stream(dev))
        return cls(nodes, leaves, T, depth, n_feat, cf, base, tree_block, dev)

    @classmethod
    def from_model(cls, model, tree_block=64):
        """Compile a fitted WarpGBM's forest (nested dicts) for serving.
        A multi-output tree (one vector of leaf values per leaf) becomes one
        tree per output, sharing its splits: output m of tree i is tree
        i * outputs + m, so output m is that forest's channel m."""
        D = model.max_depth
        slots = 2 ** (D + 1) - 1
        base = model._base_vector()
        K = len(base)
        if model.objective in ("regression", "multi_regression"):
            trees = [t for t in model.forest[: model._trees_trained] if t]
        else:
            trees = [t for rnd in model.forest[: model._trees_trained] for t in rnd]
        T = len(trees) * (K if model.objective == "multi_regression" else 1)
        per = T // max(len(trees), 1)  # compiled trees per stored tree
        st = np.zeros((T, slots), np.int32)
        ft = np.zeros((T, slots), np.int32)
        bt = np.zeros((T, slots), np.int32)
        vt = np.zeros((T, slots), np.float32)
        for i, tree in enumerate(trees):
            r = i * per
            stack = [(tree, 0)]
            while stack:
                node, h = stack.pop()
                if "leaf_value" in node:
                    st[r:r + per, h] = 2
                    vt[r:r + per, h] = np.asarray(node["leaf_value"], np.float32).reshape(-1)[:per]
                else:
                    st[r:r + per, h] = 1
                    ft[r:r + per, h] = model._feature_of(node["feature"])
                    bt[r:r + per, h] = int(node["bin"])
                    stack.append((node["left"], 2 * h + 1))
                    stack.append((node["right"], 2 * h + 2))
        return cls.from_tables(st, ft, bt, vt, D, model.learning_rate, model.num_features, base, K, tree_block,
                               model.device)

    # -------------------------------------------------------------- serving
    def workspace(self, n_rows):
        need = _lib.query("wgbm_heap_predict_workspace_bytes", n_rows, self.n_classes, self.n_blocks)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def predict_into(self, codes, out, stream=None, workspace=None):
        """out[n, K] = base + sum of trees for int8 codes [n, n_feat] (async)."""
        n = codes.shape[0]
        if codes.dtype != torch.int8 or codes.dim() != 2 or codes.shape[1] != self.n_feat:
            raise RuntimeError(f"codes must be int8 [n, {self.n_feat}]")
        ws = workspace if workspace is not None else self.workspace(n)
        st = stream if stream is not None else _stream(self.device)
        _lib.call("wgbm_heap_predict", codes.data_ptr(), n, self.n_feat, self.nodes.data_ptr(),
                  self.leaves.data_ptr(), self.n_trees, self.depth, self.tree_block, self.class_first.data_ptr(),
                  self.blocks_of_class.data_ptr(), self.n_classes, self.n_blocks, self.base.data_ptr(),
                  out.data_ptr(), ws.data_ptr(), ws.numel(), st)
        return out

    def predict_tree_order(self, codes, stream=None):
        """[n, K] scores with every class's trees added one by one in tree
        order after the base (the order training accumulates scores in)."""
        n = codes.shape[0]
        if codes.dtype != torch.int8 or codes.dim() != 2 or codes.shape[1] != self.n_feat:
            raise RuntimeError(f"codes must be int8 [n, {self.n_feat}]")
        codes = codes.contiguous()
        out = torch.empty((n, self.n_classes), dtype=torch.float32, device=self.device)
        st = stream if stream is not None else _stream(self.device)
        _lib.call("wgbm_heap_predict_tree_order", codes.data_ptr(), n, self.n_feat, self.nodes.data_ptr(),
                  self.leaves.data_ptr(), self.n_trees, self.depth, self.class_first.data_ptr(), self.n_classes,
                  self.base.data_ptr(), out.data_ptr(), st)
        return out

    def predict(self, codes):
        out = torch.empty((codes.shape[0], self.n_classes), dtype=torch.float32, device=self.device)
        return self.predict_into(codes.contiguous(), out)
