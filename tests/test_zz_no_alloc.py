"""GPU test of wgbm_heap_predict_noalloc, the serving predictor that
allocates nothing (its partial sums live in a file-scope device array).

Two calls in a row through the C ABI, the second with fewer rows (the array
keeps the first call's contents), each equal bit for bit to
wgbm_heap_predict with a workspace on the same forest and rows; an input
beyond the array's capacity is refused with an error before any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _forest(dev, K, T_per_class, depth, F, seed):
    from warpgbm_amd.serving import HeapForest
    rng = np.random.default_rng(seed)
    T = K * T_per_class
    slots = 2 ** (depth + 1) - 1
    st = np.zeros((T, slots), np.int32)
    ft = np.zeros((T, slots), np.int32)
    bt = np.zeros((T, slots), np.int32)
    vt = np.zeros((T, slots), np.float32)
    for t in range(T):  # random complete-ish trees: leaves at random depths
        stack = [(0, 0)]
        while stack:
            h, d = stack.pop()
            if d == depth or (d > 1 and rng.random() < 0.2):
                st[t, h], vt[t, h] = 2, rng.standard_normal()
            else:
                st[t, h], ft[t, h], bt[t, h] = 1, rng.integers(0, F), rng.integers(0, 31)
                stack += [(2 * h + 1, d + 1), (2 * h + 2, d + 1)]
    return HeapForest.from_tables(st, ft, bt, vt, depth, 0.05, F, rng.standard_normal(K).astype(np.float32), K,
                                  tree_block=16, device=dev)


def _noalloc(hf, codes, out, stream):
    from warpgbm_amd import _lib
    _lib.call("wgbm_heap_predict_noalloc", codes.data_ptr(), codes.shape[0], hf.n_feat, hf.nodes.data_ptr(),
              hf.leaves.data_ptr(), hf.n_trees, hf.depth, hf.tree_block, hf.class_first.data_ptr(),
              hf.blocks_of_class.data_ptr(), hf.n_classes, hf.n_blocks, hf.base.data_ptr(), out.data_ptr(), stream)


@pytest.mark.parametrize("K", [1, 3])
def test_noalloc_predict_equals_workspace_predict_twice(K):
    dev = torch.device("cuda:0")
    F = 60
    hf = _forest(dev, K, 70, 6, F, seed=K)
    rng = np.random.default_rng(7 + K)
    stream = torch.cuda.current_stream(dev).cuda_stream
    for n in (40000, 333):  # the second call is smaller: stale partials of the first stay in the array
        codes = torch.from_numpy(rng.integers(0, 32, size=(n, F), dtype=np.int8)).to(dev)
        ref = hf.predict(codes)
        out = torch.full((n, K), float("nan"), dtype=torch.float32, device=dev)
        _noalloc(hf, codes, out, stream)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"n={n}"


def test_noalloc_predict_rejects_inputs_beyond_its_array():
    from warpgbm_amd import _lib
    dev = torch.device("cuda:0")
    hf = _forest(dev, 2, 40, 4, 10, seed=5)
    n = (4 << 20) // (hf.n_classes * hf.n_blocks) + 1
    codes = torch.zeros((1, 10), dtype=torch.int8, device=dev)  # never read: the size check comes first
    out = torch.empty((1, 2), dtype=torch.float32, device=dev)
    lib = _lib.load()
    rc = lib.wgbm_heap_predict_noalloc(codes.data_ptr(), n, 10, hf.nodes.data_ptr(), hf.leaves.data_ptr(),
                                       hf.n_trees, hf.depth, hf.tree_block, hf.class_first.data_ptr(),
                                       hf.blocks_of_class.data_ptr(), hf.n_classes, hf.n_blocks, hf.base.data_ptr(),
                                       out.data_ptr(), None)
    assert rc != 0 and b"at most" in lib.wgbm_last_error()
