"""GPU parity of the reference ops and per-node helpers against the oracle.

Every test draws seeded inputs, runs the HIP library through the C ABI (via
the reference-compatible op module where one exists) and compares with the
FP64 oracle: integer/byte/index results bit-exact, floating results within
the tolerance stated in each test.
"""
import numpy as np
import pytest

from oracle import gbm_oracle as O
import _parity as PAR

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _hist_gpu(dev, codes, g, h, rows, feats, eras, E, B):
    from warpgbm_amd.cuda import node_kernel as nk
    G = torch.zeros((E, len(feats), B), dtype=torch.float32, device=dev)
    H = torch.zeros_like(G)
    nk.compute_histogram3(_t(codes, dev), _t(g.astype(np.float32), dev), _t(h.astype(np.float32), dev),
                          _t(rows.astype(np.int32), dev), _t(np.asarray(feats, np.int32), dev),
                          _t(eras.astype(np.int32), dev), G, H, B, 64, 4)
    torch.cuda.synchronize()
    return G.cpu().numpy(), H.cpu().numpy()


def _rel_err(got, ref, scale):
    return float(np.max(np.abs(got.astype(np.float64) - ref)) / max(scale, 1e-30))


def _check_hist_bins(name, Gg, Hg, codes, g, h, rows, feats, eras, E, B):
    """Every bin within 1e-5 of the FP64 oracle relative to its own sum of
    |g| (|h|); empty bins exactly zero."""
    Go, Ho = O.histogram(codes, g, h, rows, feats, eras, E, B)
    AG, AH = O.histogram(codes, np.abs(g.astype(np.float64)), np.abs(h.astype(np.float64)), rows, feats, eras, E, B)
    PAR.check_hist(name + " G", Gg, Go, AG)
    PAR.check_hist(name + " H", Hg, Ho, AH)
    return Go, Ho, AG, AH


@pytest.mark.parametrize("N,F,B,E,subset", [
    (5000, 50, 32, 1, False),
    (20000, 37, 7, 1, True),
    (30000, 100, 5, 4, False),
    (3000, 1000, 32, 1, False),
    (40000, 64, 127, 1, False),
    (60000, 6, 5, 20000, False),  # 80 KB of era counters in LDS
])
def test_histogram_matches_oracle(dev, N, F, B, E, subset):
    rng = np.random.default_rng(N + F)
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = rng.standard_normal(N).astype(np.float32)
    h = rng.uniform(0.5, 1.5, N).astype(np.float32)
    eras = rng.integers(0, E, N).astype(np.int32)
    rows = np.sort(rng.choice(N, size=N * 2 // 3, replace=False)).astype(np.int32)
    feats = rng.permutation(F)[: F // 2] if subset else np.arange(F)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, feats, eras, E, B)
    _check_hist_bins("histogram op", Gg, Hg, codes, g, h, rows, feats, eras, E, B)
    # invariants: every row lands in exactly one bin per feature
    np.testing.assert_allclose(Gg.sum(axis=(0, 2)), g[rows].astype(np.float64).sum(), rtol=1e-4, atol=1e-2)


@pytest.mark.parametrize("cfg", [(0, 8, 1, 0, 0), (0, 16, 1, 0, 0), (0, 8, 2, 0, 0), (0, 16, 2, 0, 0),
                                 (0, 8, 4, 0, 0), (0, 16, 4, 0, 256), (0, 4, 4, 16, 128), (0, 8, 8, 0, 1024),
                                 (0, 8, 16, 0, 0), (1024, 16, 4, 0, 256)])
@pytest.mark.parametrize("subset", [False, True])
def test_histogram_every_tiling_matches_oracle(dev, cfg, subset):
    """Every (chunk, waves, features per lane, depth) the autotuner can
    install builds the same histogram, for identity columns and for a
    feature subset (byte gathers)."""
    from warpgbm_amd import _lib
    rng = np.random.default_rng(17)
    N, F, B, E = 20000, 300, 8, 2
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = rng.standard_normal(N).astype(np.float32)
    h = rng.uniform(0.5, 1.5, N).astype(np.float32)
    eras = rng.integers(0, E, N).astype(np.int32)
    rows = np.sort(rng.choice(N, size=N * 3 // 4, replace=False)).astype(np.int32)
    feats = rng.permutation(F)[: F // 2] if subset else np.arange(F)
    _lib.call("wgbm_set_hist_config", *cfg)
    try:
        Gg, Hg = _hist_gpu(dev, codes, g, h, rows, feats, eras, E, B)
    finally:
        _lib.call("wgbm_set_hist_config", 0, 0, 0, 0, 0)
    _check_hist_bins("histogram op, every tiling", Gg, Hg, codes, g, h, rows, feats, eras, E, B)


@pytest.mark.parametrize("const_h", [True, False])
def test_histogram_bins_of_many_rows(dev, const_h):
    """Bins of 150k rows (2 bins, 300k rows, one era): with one hessian value
    H is the exact count * h (a fixed-point count * h product wrapped past
    2^63 from 2^16 rows); with varying hessians the per-segment fixed-point
    sum stays within 64 bits (per-row bits follow the matrix's row count)."""
    rng = np.random.default_rng(77)
    N, F, B = 300_000, 3, 2
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = rng.standard_normal(N).astype(np.float32)
    h = np.full(N, 0.75, np.float32) if const_h else rng.uniform(0.5, 1.0, N).astype(np.float32)
    eras = np.zeros(N, np.int32)
    rows = np.arange(N, dtype=np.int32)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, 1, B)
    Go, Ho, _, _ = _check_hist_bins("histogram op, 150k-row bins", Gg, Hg, codes, g, h, rows, np.arange(F), eras, 1, B)
    if const_h:
        np.testing.assert_array_equal(Hg, Ho.astype(np.float32))  # count * 0.75, exact


def _hdr_gradients(rng, n, codes_col, B, lo=-5.0, hi=3.0):
    """Gradients whose magnitude follows one feature's bin: bin b holds
    |g| = 10^(lo + (hi - lo) b / (B - 1)) (times a factor in [1, 2)), random
    signs, so the smallest bins hold only rows 10^(hi - lo) below the
    node's largest |g| -- the case a fixed-point scale taken from the node's
    max |g| rounds worst."""
    mag = 10.0 ** (lo + (hi - lo) * codes_col.astype(np.float64) / (B - 1))
    return (mag * rng.uniform(1.0, 2.0, n) * rng.choice([-1.0, 1.0], n)).astype(np.float32)


@pytest.mark.parametrize("hess", ["constant", "varying"])
@pytest.mark.parametrize("E", [1, 3])
def test_histogram_high_dynamic_range(dev, hess, E):
    """|g| from 1e-5 to 2e3 in one node, small magnitudes concentrated in
    whole bins.  Constant hessians take the count-packed 34-bit words with
    the exact fine path for small rows; varying hessians the 49-bit words.
    Every bin of every feature stays within 1e-5 of its own sum of |g|."""
    rng = np.random.default_rng(101 + E)
    N, F, B = 40000, 24, 16
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = _hdr_gradients(rng, N, codes[:, 0], B)
    h = np.ones(N, np.float32) if hess == "constant" else rng.uniform(0.1, 3.0, N).astype(np.float32)
    eras = rng.integers(0, E, N).astype(np.int32)
    rows = np.sort(rng.choice(N, size=N * 3 // 4, replace=False)).astype(np.int32)
    feats = np.arange(F)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, feats, eras, E, B)
    _check_hist_bins(f"histogram op, |g| 1e-5..2e3 ({hess} h)", Gg, Hg, codes, g, h, rows, feats, eras, E, B)
    # the small bins really are 1e-8 of the node's max
    assert np.abs(g[rows][codes[rows, 0] == 0]).max() < 2e-8 * np.abs(g[rows]).max()
    # a second call on the same op gives the same bits (fine words left at zero)
    Gg2, Hg2 = _hist_gpu(dev, codes, g, h, rows, feats, eras, E, B)
    assert np.array_equal(Gg, Gg2) and np.array_equal(Hg, Hg2)


def test_histogram_late_round_multiclass_node(dev):
    """A late-round softmax node: most rows classified with probability
    within 1e-6..1e-3 of 0 or 1 (|g| tiny, hessians clamped near 1e-6), a
    few misclassified rows with |g| near 1; confident rows cluster by
    feature value.  Gradients and hessians of one class, as the objective
    computes them in float32."""
    rng = np.random.default_rng(55)
    N, F, B, K = 30000, 16, 32, 5
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    y = rng.integers(0, K, N)
    margin = 2.0 + 14.0 * (codes[:, 0].astype(np.float64) / (B - 1))  # confident where feature 0 is high
    scores = rng.standard_normal((N, K)) * 0.5
    scores[np.arange(N), y] += margin
    wrong = rng.random(N) < 0.02
    scores[wrong, (y[wrong] + 1) % K] += 2 * margin[wrong]
    gk, hk = O.softmax_grad_hess(scores.astype(np.float32), y, np.ones(N))
    g = gk[:, 2].astype(np.float32)
    h = hk[:, 2].astype(np.float32)
    rows = np.arange(N, dtype=np.int32)
    eras = np.zeros(N, np.int32)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, 1, B)
    nz = np.abs(g[g != 0])
    assert nz.min() < 1e-6 * nz.max()  # a high dynamic range really is present
    _check_hist_bins("histogram op, late-round softmax node", Gg, Hg, codes, g, h, rows, np.arange(F), eras, 1, B)


def test_histogram_counts_exact_and_drops_out_of_range(dev):
    rng = np.random.default_rng(3)
    N, F, B = 10000, 20, 16
    codes = rng.integers(-3, B + 3, size=(N, F)).astype(np.int8)
    g = np.ones(N, np.float32)
    h = np.ones(N, np.float32)
    rows = np.arange(N, dtype=np.int32)
    eras = np.zeros(N, np.int32)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, 1, B)
    Go, Ho = O.histogram(codes, g, h, rows, np.arange(F), eras, 1, B)
    assert np.array_equal(Hg, Ho.astype(np.float32))  # integer counts are exact in FP32
    assert np.array_equal(Gg, Go.astype(np.float32))


def test_histogram_deterministic_and_era_order_independent(dev):
    rng = np.random.default_rng(11)
    N, F, B, E = 50000, 40, 32, 5
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = rng.standard_normal(N).astype(np.float32)
    h = np.ones(N, np.float32)
    eras = rng.integers(0, E, N).astype(np.int32)
    rows = np.arange(N, dtype=np.int32)
    a = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, E, B)
    b = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, E, B)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the op sorts rows by era stably: an era-grouped list gives the same bits
    grouped = np.argsort(eras, kind="stable").astype(np.int32)
    c = _hist_gpu(dev, codes, g, h, grouped, np.arange(F), eras, E, B)
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_histogram_empty_rows(dev):
    codes = np.zeros((10, 4), np.int8)
    G, H = _hist_gpu(dev, codes, np.ones(10, np.float32), np.ones(10, np.float32), np.zeros(0, np.int32),
                     np.arange(4), np.zeros(10, np.int32), 1, 8)
    assert not G.any() and not H.any()


@pytest.mark.parametrize("E", [1, 3])
def test_compute_split_and_select_match_oracle(dev, E):
    from warpgbm_amd.cuda import node_kernel as nk
    from warpgbm_amd import _lib
    rng = np.random.default_rng(5 + E)
    F, B = 30, 16
    N = 4000
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    y = (codes[:, 3] > 7).astype(np.float32) * 2.0 - (codes[:, 5] > 3) + 0.1 * rng.standard_normal(N)
    g = (y - y.mean()).astype(np.float32)
    h = np.ones(N, np.float32)
    eras = (np.arange(N) * E // N).astype(np.int32)
    rows = np.arange(N, dtype=np.int32)
    Gg, Hg = _hist_gpu(dev, codes, g, h, rows, np.arange(F), eras, E, B)
    Gt, Ht = _t(Gg, dev), _t(Hg, dev)
    gain = torch.empty((E, F, B - 1), dtype=torch.float32, device=dev)
    d = torch.empty_like(gain)
    nk.compute_split(Gt, Ht, 0.0, 20.0, 1e-6, gain, d, 64)
    out = torch.empty(2, dtype=torch.int32, device=dev)
    _lib.call("wgbm_select_split", gain.data_ptr(), d.data_ptr(), E, F, B, 0.0, None, None, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    Go, Ho = O.histogram(codes, g, h, rows, np.arange(F), eras, E, B)
    AG, AH = O.histogram(codes, np.abs(g.astype(np.float64)), h, rows, np.arange(F), eras, E, B)
    gain_o, dir_o = O.split_gains(Go, Ho, 20.0, 1e-6)
    gg = gain.cpu().numpy()
    # per threshold: the error the per-bin histogram bound can propagate
    gbound, dbound, margin = PAR.gain_error_bound(Go, Ho, AG, AH, 1e-6)
    ok = (O.split_gains(Go, Ho, 20.0, 1e-6)[1] != 0)  # thresholds past min_child_weight
    err = np.abs(gg.astype(np.float64) - gain_o)
    assert np.all(err[ok] <= gbound[ok]), float(np.max((err / gbound)[ok]))
    assert np.all(gg[~ok] == 0.0)
    PAR.record("compute_split gain (error / per-threshold bound)", float(np.max((err / gbound)[ok])), 1.0)
    dd = d.cpu().numpy()
    clear = ok & (margin > dbound)  # direction fixed by the FP64 values beyond any histogram error
    assert clear.sum() > 0.9 * ok.sum()
    assert np.array_equal(dd[clear], dir_o[clear])
    f, b = O.select_split(gain_o, dir_o, 0.0)
    assert tuple(out.cpu().numpy().tolist()) == (f, b)


def test_select_conflicting_eras_is_leaf(dev):
    """Two eras with opposite monotone signal and a +1 constraint -> no split
    (mirrors warpgbm tests/test_monotonic_constraints.py:124-149)."""
    from warpgbm_amd import _lib
    E, F, B = 2, 1, 8
    gain = np.zeros((E, F, B - 1), np.float32) + 1.0
    d = np.zeros((E, F, B - 1), np.float32)
    d[0] = 1.0
    d[1] = -1.0
    req = np.array([-1], np.int8)  # monotone increasing -> required direction -1
    out = torch.empty(2, dtype=torch.int32, device=dev)
    gd, dd, rd = _t(gain, dev), _t(d, dev), _t(req, dev)  # held until the kernel has run
    _lib.call("wgbm_select_split", gd.data_ptr(), dd.data_ptr(), E, F, B, 0.0, rd.data_ptr(), None, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert out.cpu().tolist() == [-1, -1]
    assert O.select_split(gain, d, 0.0, req) == (-1, -1)


def test_partition_rows_is_stable_and_exact(dev):
    from warpgbm_amd import _lib
    rng = np.random.default_rng(9)
    N, F = 100000, 12
    codes = rng.integers(0, 32, size=(N, F), dtype=np.int8)
    rows = np.sort(rng.choice(N, 70001, replace=False)).astype(np.int32)
    ws = torch.empty(_lib.query("wgbm_partition_workspace_bytes", rows.size), dtype=torch.uint8, device=dev)
    out = torch.empty(rows.size, dtype=torch.int32, device=dev)
    nl = torch.empty(1, dtype=torch.int32, device=dev)
    ct, rt = _t(codes, dev), _t(rows, dev)
    _lib.call("wgbm_partition_rows", ct.data_ptr(), N, F, rt.data_ptr(), rows.size, 7, 13, out.data_ptr(),
              nl.data_ptr(), ws.data_ptr(), ws.numel(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    m = codes[rows, 7] <= 13
    ref = np.concatenate([rows[m], rows[~m]])
    assert int(nl.item()) == int(m.sum())
    assert np.array_equal(out.cpu().numpy(), ref)


def test_leaf_value_and_score_update(dev):
    from warpgbm_amd import _lib
    rng = np.random.default_rng(2)
    N = 30000
    g = rng.standard_normal(N).astype(np.float32)
    h = rng.uniform(0.1, 2, N).astype(np.float32)
    rows = np.sort(rng.choice(N, 12345, replace=False)).astype(np.int32)
    pred = rng.standard_normal(N).astype(np.float32)
    ws = torch.empty(_lib.query("wgbm_leaf_workspace_bytes", rows.size), dtype=torch.uint8, device=dev)
    out = torch.empty(1, dtype=torch.float32, device=dev)
    pt = _t(pred, dev)
    gt, ht, rt = _t(g, dev), _t(h, dev), _t(rows, dev)
    _lib.call("wgbm_leaf_value", gt.data_ptr(), ht.data_ptr(), rt.data_ptr(), rows.size, N, 1e-6, 0.1,
              pt.data_ptr(), 1, 0, out.data_ptr(), None, ws.data_ptr(), ws.numel(),
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ref = O.leaf_value(g, h, rows, 1e-6)
    v = float(out.item())
    assert abs(v - ref) <= 1e-6 * max(abs(ref), 1e-6)
    exp = pred.copy()
    exp[rows] = exp[rows] + np.float32(np.float32(0.1) * np.float32(v))
    assert np.array_equal(pt.cpu().numpy(), exp)


@pytest.mark.parametrize("N,B", [(20000, 32), (12345, 10), (1000, 127), (7, 5)])
def test_quantile_edges_and_binning_bit_exact(dev, N, B):
    from warpgbm_amd import _lib
    rng = np.random.default_rng(N)
    F = 8
    X = rng.standard_normal((N, F)).astype(np.float32)
    X[:, 3] = np.round(X[:, 3])  # heavy ties
    X[5 % N, 6] = np.nan         # NaN column -> NaN edges, codes 0
    Xt = _t(X, dev)
    srt = torch.sort(Xt, dim=0).values.contiguous()
    edges = torch.empty((F, B - 1), dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("wgbm_edges_from_sorted", srt.data_ptr(), N, F, B, edges.data_ptr(), B - 1, st)
    codes = torch.empty((N, F), dtype=torch.int8, device=dev)
    ws = torch.empty(_lib.query("wgbm_bin_matrix_workspace_bytes", F, B - 1), dtype=torch.uint8, device=dev)
    _lib.call("wgbm_bin_matrix", Xt.data_ptr(), N, F, edges.data_ptr(), B - 1, codes.data_ptr(), ws.data_ptr(),
              ws.numel(), st)
    torch.cuda.synchronize()
    ref_codes, ref_edges = O.bin_matrix(X, B)
    e = edges.cpu().numpy()
    assert np.array_equal(np.isnan(e), np.isnan(ref_edges))
    ok = ~np.isnan(e)
    # equal values, equal bits except the sign of a zero edge (the device
    # sort may order -0.0 and +0.0 either way; codes cannot change)
    assert np.array_equal(e[ok], ref_edges[ok])
    nz = ok & (e != 0)
    assert np.array_equal(e[nz].view(np.int32), ref_edges[nz].view(np.int32))
    assert np.array_equal(codes.cpu().numpy(), ref_codes)


@pytest.mark.parametrize("N,F,nb", [(3000, 100, 31), (777, 68, 127), (513, 7, 9), (64, 4, 1)])
def test_bin_matrix_any_edges(dev, N, F, nb):
    """Whole-matrix binning against the early-exit scan for edges that are
    not sorted and hold NaN (tile path: F % 4 == 0; scalar path otherwise)."""
    from warpgbm_amd import _lib
    rng = np.random.default_rng(F + nb)
    X = rng.standard_normal((N, F)).astype(np.float32)
    X[rng.random((N, F)) < 0.01] = np.nan
    edges = np.sort(rng.standard_normal((F, nb)), axis=1).astype(np.float32)
    edges[::3] = rng.standard_normal((len(edges[::3]), nb)).astype(np.float32)  # unsorted rows
    edges[1::5, nb // 2] = np.nan
    codes = torch.empty((N, F), dtype=torch.int8, device=dev)
    ws = torch.empty(_lib.query("wgbm_bin_matrix_workspace_bytes", F, nb), dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    Xd, Ed = _t(X, dev), _t(edges, dev)  # held: a freed temporary's memory could be reused by the next upload
    _lib.call("wgbm_bin_matrix", Xd.data_ptr(), N, F, Ed.data_ptr(), nb, codes.data_ptr(), ws.data_ptr(), ws.numel(),
              st)
    torch.cuda.synchronize()
    assert np.array_equal(codes.cpu().numpy(), O.bin_with_edges(X, edges))


@pytest.mark.parametrize("N,F,B", [(20000, 9, 32), (12345, 5, 10), (1000, 3, 127), (7, 4, 5), (1, 2, 32),
                                   (2, 3, 3), (100003, 2, 64)])
def test_quantile_select_edges_bit_exact(dev, N, F, B):
    """Radix-select edges equal the oracle's sorted-column edges bit for bit
    (both order -0.0 before +0.0), ties, signed zeros and NaN columns included."""
    from warpgbm_amd import _lib
    rng = np.random.default_rng(N + B)
    X = rng.standard_normal((N, F)).astype(np.float32)
    if F > 2:
        X[:, 1] = np.round(X[:, 1] * 2)  # heavy ties
        X[:, 2] = np.where(rng.random(N) < 0.5, np.float32(0.0), np.float32(-0.0))  # signed zeros only
    if N > 5 and F > 3:
        X[5, 3] = np.nan
    ld = F + 3  # a row stride wider than the columns
    Xp = np.zeros((N, ld), np.float32)
    Xp[:, :F] = X
    edges = torch.full((F, B + 1), 7.0, dtype=torch.float32, device=dev)
    ws = torch.empty(_lib.query("wgbm_quantile_edges_workspace_bytes", N, F), dtype=torch.uint8, device=dev)
    Xd = _t(Xp, dev)
    _lib.call("wgbm_quantile_edges", Xd.data_ptr(), N, F, ld, B, edges.data_ptr(), B + 1, ws.data_ptr(), ws.numel(),
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    e = edges.cpu().numpy()
    ref = np.stack([O.quantile_edges_f32(X[:, f], B) for f in range(F)])
    assert np.array_equal(e[:, :B - 1].view(np.int32), ref.view(np.int32)) or np.array_equal(
        np.isnan(e[:, :B - 1]), np.isnan(ref)) and np.array_equal(e[:, :B - 1][~np.isnan(ref)].view(np.int32),
                                                                  ref[~np.isnan(ref)].view(np.int32))
    assert np.all(e[:, B - 1:] == 7.0)  # padding of the edge rows untouched


def test_bin_column_reference_op(dev):
    from warpgbm_amd.cuda import node_kernel as nk
    rng = np.random.default_rng(1)
    x = rng.standard_normal(5001).astype(np.float32)
    x[7] = np.nan
    edges = np.sort(rng.standard_normal(9)).astype(np.float32)
    out = torch.empty(x.size, dtype=torch.int8, device=dev)
    nk.custom_cuda_binner(_t(x, dev), _t(edges, dev), out)
    torch.cuda.synchronize()
    ref = O.bin_codes(x, edges)
    assert np.array_equal(out.cpu().numpy(), ref)
    assert np.array_equal(ref[~np.isnan(x)], np.searchsorted(edges, x[~np.isnan(x)], side="right"))


def test_predict_forest_reference_op(dev):
    from warpgbm_amd.cuda import node_kernel as nk
    rng = np.random.default_rng(4)
    N, F, B = 3000, 10, 16
    codes = rng.integers(0, B, size=(N, F), dtype=np.int8)
    g = rng.standard_normal(N)
    h = np.ones(N)
    trees = []
    for t in range(7):
        rec = O.grow_tree(codes, g * (t + 1), h, np.zeros(N, np.int32), 1, B, 4, 20, 0.0, 1e-6)
        trees.append(O.flatten_tree(rec.tree, 2 ** 5))
    flat = np.stack(trees).astype(np.float32)
    out = torch.full((N,), 0.5, dtype=torch.float32, device=dev)
    nk.predict_forest(_t(codes, dev), _t(flat, dev), 0.1, out)
    torch.cuda.synchronize()
    ref = O.predict_forest_flat(codes, flat, 0.1, base=0.5)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)
    # chunk invariance (warpgbm tests/test_predict_chunking.py:19-40), bitwise here
    o1 = torch.full((N // 2,), 0.5, dtype=torch.float32, device=dev)
    nk.predict_forest(_t(codes[: N // 2], dev), _t(flat, dev), 0.1, o1)
    torch.cuda.synchronize()
    assert np.array_equal(o1.cpu().numpy(), out.cpu().numpy()[: N // 2])
