"""GPU parity: libcrest_hip.so (through the C ABI via crest_amd._native) against the
oracle's restatement of the reference on the same seeded inputs.

Bars: the lazy transform and the projection are bit-exact; FP64 sums whose order
differs (gene stats, FP64 Gram) agree to 1e-12 relative; FP32 / split-BF16 Grams to
the stated relative tolerance; eigenvalues to 1e-9 relative; kNN index sets equal
except documented distance ties.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import _native as N  # noqa: E402
from crest_amd import synth  # noqa: E402


@pytest.fixture(scope="module")
def counts():
    indptr, indices, data, _ = synth.make_counts(1500, 900, 6, density=0.07, seed=3)
    return indptr, indices, data, 900


def _frame(counts, chunk_nnz=None, drop=True):
    indptr, indices, data, g = counts
    bf = ca.BioFrame.from_csr(indptr, indices, data, g)
    if chunk_nnz:
        bf.chunk_nnz = chunk_nnz
    if drop:  # filtered cells and genes exercise the maps
        rng = np.random.default_rng(0)
        bf = bf.filter_cells(rng.random(bf.n_obs) > 0.05).filter_genes(rng.random(bf.n_vars) > 0.1)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    return bf


def test_materialize_hand_values():
    """kernels.rs:564-578: normalised values 25/75/100, CSR == COO."""
    cells = np.array([0, 0, 1, 2, 2], np.uint32)
    genes = np.array([0, 1, 0, 1, 2], np.uint32)
    vals = np.array([2, 6, 5, 1, 3], np.float32)
    cm = np.array([0, -1, 1], np.int64)
    gm = np.array([0, 1, -1], np.int32)
    r, g, x = N.materialize(genes, vals, cm, gm, 100.0, False, cells, None, 0)
    assert r.tolist() == [0, 0, 1] and g.tolist() == [0, 1, 1]
    assert x.tolist() == [25.0, 75.0, 100.0]
    indptr = np.array([0, 2, 3, 5], np.int64)
    r2, g2, x2 = N.materialize(genes, vals, cm, gm, 100.0, False, None, indptr, 0)
    assert r2.tolist() == r.tolist() and g2.tolist() == g.tolist() and x2.tolist() == x.tolist()


@pytest.mark.parametrize("chunk_nnz", [None, 3000])
def test_materialize_bit_exact(counts, oracle, chunk_nnz):
    bf = _frame(counts, chunk_nnz)
    for ctx in bf.iter_ctx():
        got = N.materialize(*ctx)
        ref = oracle.materialize(*ctx)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_gene_stats(counts, oracle):
    bf = _frame(counts, 4000)
    got = np.zeros((bf.n_vars, 5))
    ref = np.zeros((bf.n_vars, 5))
    for ctx in bf.iter_ctx():
        N.gene_stats(*ctx, got)
        oracle.gene_stats(*ctx, ref)
    np.testing.assert_array_equal(got[:, 4], ref[:, 4])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("chunk_nnz", [None, 5000])
def test_qc_bit_exact(counts, oracle, chunk_nnz):
    """kernels.rs:162-195: integer counts, so every sum is exact and the GPU result
    equals the oracle's bit for bit (with filtered cells/genes and chunking)."""
    bf = _frame(counts, chunk_nnz)
    rng = np.random.default_rng(1)
    flag = rng.random(bf.n_vars) < 0.05
    got = [np.zeros(bf.n_obs), np.zeros(bf.n_obs, np.uint32), np.zeros(bf.n_obs), np.zeros(bf.n_vars, np.uint32),
           np.zeros(bf.n_vars)]
    ref = [a.copy() for a in got]
    for ctx in bf.iter_ctx(transform=False):
        N.qc(*ctx, flag, *got)
        oracle.qc(*ctx, flag, *ref)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    assert got[0].sum() > 0 and got[3].sum() > 0


def test_calculate_qc_metrics(counts):
    """pp.py:25-47 against dense numpy on the kept cells and genes."""
    indptr, indices, data, g = counts
    bf = ca.BioFrame.from_csr(indptr, indices, data, g)
    rng = np.random.default_rng(2)
    bf = bf.filter_cells(rng.random(bf.n_obs) > 0.1).filter_genes(rng.random(bf.n_vars) > 0.2)
    bf.var["gene_name"] = np.array([("MT-%d" % i) if i % 11 == 0 else ("G%d" % i) for i in range(bf.n_vars)])
    obs, var = ca.pp.calculate_qc_metrics(bf, inplace=False)
    D = bf.to_dense(transform=False).astype(np.float64)
    mt = np.array([n.startswith("MT-") for n in bf.var_names])
    np.testing.assert_array_equal(obs["total_counts"], D.sum(1))
    np.testing.assert_array_equal(obs["n_genes_by_counts"], (D != 0).sum(1))
    pct = np.divide(D[:, mt].sum(1), D.sum(1), out=np.zeros(len(D)), where=D.sum(1) > 0) * 100.0
    np.testing.assert_array_equal(obs["pct_counts_mt"], pct)
    np.testing.assert_array_equal(var["n_cells_by_counts"], (D != 0).sum(0))
    np.testing.assert_array_equal(var["total_counts"], D.sum(0))
    with pytest.raises(ValueError):
        ctx = next(bf.iter_ctx(transform=False))
        N.qc(*ctx, np.zeros(3, bool), np.zeros(bf.n_obs), np.zeros(bf.n_obs, np.uint32), np.zeros(bf.n_obs),
             np.zeros(bf.n_vars, np.uint32), np.zeros(bf.n_vars))


def _gram_inputs(bf, oracle, d_top=300):
    st = np.zeros((bf.n_vars, 5))
    for ctx in bf.iter_ctx():
        oracle.gene_stats(*ctx, st)
    n = bf.n_obs
    mean = st[:, 0] / n
    var = np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0)
    sd = np.sqrt(var)
    sd[sd == 0] = 1
    mask = np.zeros(bf.n_vars, bool)
    mask[np.argsort(-var)[:d_top]] = True
    d = int(mask.sum())
    gmap = np.full(bf.n_vars, -1, np.int32)
    gmap[mask] = np.arange(d, dtype=np.int32)
    return gmap, np.ascontiguousarray(mean[mask]), np.ascontiguousarray(sd[mask]), d


@pytest.mark.parametrize("prec,rtol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-5), ("fp16", 4e-6)])
def test_gram_accumulate(counts, oracle, prec, rtol):
    bf = _frame(counts, 5000)
    gmap, mean, sd, d = _gram_inputs(bf, oracle)
    ca.set_precision(prec)
    try:
        g_got, c_got = np.zeros((d, d)), np.zeros(d)
        g_ref, c_ref = np.zeros((d, d)), np.zeros(d)
        for ctx in bf.iter_ctx():
            N.gram_accumulate(*ctx, gmap, g_got, c_got, True, mean, sd, 10.0)
            oracle.gram_accumulate(*ctx, gmap, g_ref, c_ref, True, mean, sd, 10.0)
    finally:
        ca.set_precision("fp64")
    iu = np.triu_indices(d)
    scale = np.abs(g_ref[iu]).max()
    assert np.abs(g_got[iu] - g_ref[iu]).max() <= rtol * scale
    assert np.all(np.tril(g_got, -1) == 0)  # lower triangle untouched
    np.testing.assert_allclose(c_got, c_ref, rtol=1e-12, atol=1e-9)


def test_fp16_gram_needs_clipped_values(counts, oracle):
    """Split-FP16 tiles hold |s| <= 65504, which clipping at max_value guarantees."""
    bf = _frame(counts, 5000)
    gmap, mean, sd, d = _gram_inputs(bf, oracle)
    ca.set_precision("fp16")
    try:
        ctx = next(bf.iter_ctx())
        with pytest.raises(ValueError):
            N.gram_accumulate(*ctx, gmap, np.zeros((d, d)), np.zeros(d), False, None, None, 0.0)
        with pytest.raises(ValueError):
            N.gram_accumulate(*ctx, gmap, np.zeros((d, d)), np.zeros(d), True, mean, sd, 0.0)
    finally:
        ca.set_precision("fp64")


def test_pca_from_gram(counts, oracle):
    bf = _frame(counts, 5000)
    gmap, mean, sd, d = _gram_inputs(bf, oracle, 200)
    g, c = np.zeros((d, d)), np.zeros(d)
    for ctx in bf.iter_ctx():
        oracle.gram_accumulate(*ctx, gmap, g, c, True, mean, sd, 10.0)
    var_g, load_g, tot_g = N.pca_from_gram(g, c, bf.n_obs, 20)
    var_o, load_o, tot_o = oracle.pca_from_gram(g, c, bf.n_obs, 20)
    np.testing.assert_allclose(var_g, var_o, rtol=1e-9)
    assert abs(tot_g - tot_o) <= 1e-12 * abs(tot_o)
    # per-component vectors where the eigenvalue gap allows: same sign rule
    gaps = np.minimum(np.abs(np.diff(np.r_[np.inf, var_o])), np.abs(np.diff(np.r_[var_o, 0.0])))
    ok = gaps > 1e-3 * var_o[0]
    for c_ in np.nonzero(ok)[0]:
        assert np.abs(load_g[:, c_] - load_o[:, c_]).max() < 1e-6
    assert oracle.principal_angle(load_g, load_o) < 1e-6


def test_pca_from_gram_errors():
    with pytest.raises(RuntimeError):
        N.pca_from_gram(np.zeros((3, 3)), np.zeros(3), 1, 2)
    with pytest.raises(ValueError):
        N.pca_from_gram(np.zeros((3, 2)), np.zeros(3), 10, 2)


def test_project_bit_exact(counts, oracle):
    bf = _frame(counts, 3000)
    gmap, mean, sd, d = _gram_inputs(bf, oracle, 150)
    rng = np.random.default_rng(1)
    V = rng.normal(size=(d, 17))
    out_g = np.full((bf.n_obs, 17), 0.25, np.float32)
    out_o = out_g.copy()
    for ctx in bf.iter_ctx():
        N.project(*ctx, gmap, V, out_g, True, mean, sd, 10.0)
        oracle.project(*ctx, gmap, V, out_o, True, mean, sd, 10.0)
    np.testing.assert_array_equal(out_g, out_o)


def _clustered(n, d, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-10, 10, (20, d))
    return (centers[np.arange(n) % 20] + rng.uniform(-1.5, 1.5, (n, d))).astype(np.float32)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "fp16"])
@pytest.mark.parametrize("n,d,k,scale,offset", [(600, 7, 10, 1.0, 0.0), (5000, 50, 15, 1.0, 0.0),
                                                (2049, 30, 32, 1.0, 0.0), (300, 3, 1, 1.0, 0.0),
                                                (1500, 200, 15, 1.0, 0.0), (2000, 50, 15, 3e4, 0.0),
                                                (2000, 20, 15, 1e-3, 50.0), (1000, 12, 64, 1.0, 0.0),
                                                (1200, 512, 15, 1.0, 0.0), (900, 300, 64, 1.0, 0.0)])
def test_knn_graph(oracle, prec, n, d, k, scale, offset):
    """Any width (knn.rs:337-369): rows wider than the register screens go through the
    GEMM screen and the same exact rerank."""
    x = (_clustered(n, d, n + d) * np.float32(scale) + np.float32(offset)).astype(np.float32)
    ca.set_knn_precision(prec)
    try:
        assert ca.get_knn_precision() == prec
        idx, dist = N.knn_graph(x, k)
    finally:
        ca.set_knn_precision(None)
    assert idx.shape == (n, k) and dist.shape == (n, k)
    assert not (idx == np.arange(n)[:, None]).any()
    assert (np.diff(dist, axis=1) >= 0).all()
    ref_i, ref_d2 = oracle.knn_exact_f64(x, k)
    bad, ties = oracle.knn_mismatch_ties(idx, x, np.arange(n), ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    # distances are the reference's f32 formula on the reported pairs
    pd = oracle.pair_dist(x, np.repeat(np.arange(n), k), idx.ravel()).reshape(n, k)
    np.testing.assert_array_equal(dist, pd)


@pytest.mark.parametrize("n,d,k", [(3000, 30, 100), (2500, 10, 300), (700, 5, 650), (1200, 50, 1190),
                                   (20000, 30, 212)])
def test_knn_graph_large_k(oracle, n, d, k):
    """k beyond the register lists (Scrublet's k_adj ~ 1.5 sqrt(n), doublets.py:131-137):
    same bar as the small-k path -- sets equal to the FP64 exact result up to ties,
    ascending, self excluded, distances the reference's f32 formula."""
    x = _clustered(n, d, 3 * n + d)
    idx, dist = N.knn_graph(x, k)
    assert idx.shape == (n, k) and dist.shape == (n, k)
    assert not (idx == np.arange(n)[:, None]).any()
    assert (np.diff(dist, axis=1) >= 0).all()
    assert all(len(set(r.tolist())) == k for r in idx[:: max(1, n // 200)])
    rows = np.arange(n) if n <= 3000 else np.arange(0, n, 17)
    ref_i, ref_d2 = oracle.knn_exact_f64(x, k, queries=rows)
    bad, ties = oracle.knn_mismatch_ties(idx[rows], x, rows, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(x, np.repeat(rows, k), idx[rows].ravel()).reshape(len(rows), k)
    np.testing.assert_array_equal(dist[rows], pd)


def test_knn_large_k_duplicates(oracle):
    """Many exact duplicates put ties at the selection threshold: the lowest indices win."""
    base = _clustered(300, 6, 11)
    x = np.repeat(base, 5, axis=0)  # every point five times
    k = 120
    idx, dist = N.knn_graph(x, k)
    n = len(x)
    X64 = x.astype(np.float64)
    D = ((X64[:, None, :] - X64[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    ref_i = np.lexsort((np.broadcast_to(np.arange(n), D.shape), D), axis=1)[:, :k]  # (distance, index)
    np.testing.assert_array_equal(np.sort(idx, 1), np.sort(ref_i, 1))
    pd = oracle.pair_dist(x, np.repeat(np.arange(len(x)), k), idx.ravel()).reshape(len(x), k)
    np.testing.assert_array_equal(dist, pd)


def test_knn_large_k_overflow_ties():
    """20,000 rows that are 5 distinct points: thousands of candidates share the k-th
    key, so the bounded copy overflows and the whole row is selected; ties go to the
    lowest indices."""
    base = np.random.default_rng(3).normal(size=(5, 8)).astype(np.float32) * 4
    n, k = 20000, 100
    x = base[np.arange(n) % 5]
    idx, dist = N.knn_graph(x, k)
    for i in (0, 1, 7, 12345, n - 1):
        want = [j for j in range(i % 5, n, 5) if j != i][:k]
        assert idx[i].tolist() == want
    assert (dist == 0).all()


def test_knn_query_large_k(oracle):
    R = _clustered(4000, 20, 5)
    Q = _clustered(500, 20, 6)
    k = 200
    idx, dist = N.knn_query(R, Q, k)
    assert idx.shape == (500, k)
    assert (np.diff(dist, axis=1) >= 0).all()
    ref_i, ref_d2 = oracle.knn_query_exact_f64(R, Q, k)
    bad, ties = oracle.knn_query_mismatch_ties(idx, R, Q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    X = np.vstack([Q, R])
    pd = oracle.pair_dist(X, np.repeat(np.arange(500), k), 500 + idx.ravel()).reshape(500, k)
    np.testing.assert_array_equal(dist, pd)
    idx, dist = N.knn_query(R[:150], Q[:7], 400)  # k above n_ref: every reference row
    assert idx.shape == (7, 150)
    ref_i, _ = oracle.knn_query_exact_f64(R[:150], Q[:7], 150)
    np.testing.assert_array_equal(idx, ref_i)


def test_knn_large_k_limit():
    with pytest.raises(ValueError):
        N.knn_graph(np.random.default_rng(0).normal(size=(5000, 4)).astype(np.float32), 4000)


def test_knn_precision_knob():
    ca.set_precision("bf16x3")
    try:
        assert ca.get_knn_precision() == "bf16x3"  # follows set_precision by default
        ca.set_knn_precision("fp16")
        assert ca.get_knn_precision() == "fp16" and ca.get_precision() == "bf16x3"
        with pytest.raises(ValueError):
            ca.set_knn_precision(9)
        with pytest.raises(ValueError):
            ca.set_precision(7)
        ca.set_precision("fp16")  # split-FP16 Gram + FP16 screen
        assert ca.get_precision() == "fp16"
    finally:
        ca.set_knn_precision(None)
        ca.set_precision("fp64")
    assert ca.get_knn_precision() == "fp64"


def test_knn_wide_ties_by_index():
    """300 identical columns-wide rows: every distance is 0, so each row's neighbours
    are the lowest other indices (knn.rs:68-87 keeps the earlier candidate on ties)."""
    ca.set_knn_precision("fp16")
    try:
        idx, dist = N.knn_graph(np.zeros((10, 300), np.float32), 3)
    finally:
        ca.set_knn_precision(None)
    want = np.array([[j for j in range(10) if j != i][:3] for i in range(10)])
    np.testing.assert_array_equal(idx, want)
    assert (dist == 0).all()


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "fp16"])
def test_knn_query_exact(prec):
    """tests/test_crest.py:504-511: indices equal numpy argsort, distances rtol 1e-4."""
    rng = np.random.default_rng(0)
    R = rng.normal(size=(3000, 10)).astype(np.float32)
    Q = rng.normal(size=(200, 10)).astype(np.float32)
    ca.set_knn_precision(prec)
    try:
        idx, dist = N.knn_query(R, Q, 5)
    finally:
        ca.set_knn_precision(None)
    D = ((Q[:, None, :] - R[None, :, :]) ** 2).sum(-1)
    np.testing.assert_array_equal(idx, np.argsort(D, axis=1)[:, :5])
    np.testing.assert_allclose(dist, np.sqrt(np.sort(D, axis=1)[:, :5]), rtol=1e-4)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "fp16"])
@pytest.mark.parametrize("nr,nq,d,k,shift", [(5000, 3000, 50, 15, 0.0), (700, 1300, 7, 30, 0.0),
                                             (40, 600, 3, 64, 0.0), (2000, 900, 20, 15, 4.0),
                                             (3000, 500, 400, 15, 0.0)])
def test_knn_query_clustered(oracle, prec, nr, nq, d, k, shift):
    R = _clustered(nr, d, nr + d)
    Q = (_clustered(nq, d, nq + 7 * d) + np.float32(shift)).astype(np.float32)  # shift: query set offset
    ca.set_knn_precision(prec)
    try:
        idx, dist = N.knn_query(R, Q, k)
    finally:
        ca.set_knn_precision(None)
    kk = min(k, nr)
    assert idx.shape == (nq, kk) and dist.shape == (nq, kk)
    assert (np.diff(dist, axis=1) >= 0).all()
    ref_i, ref_d2 = oracle.knn_query_exact_f64(R, Q, kk)
    bad, ties = oracle.knn_query_mismatch_ties(idx, R, Q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    X = np.vstack([Q, R])
    pd = oracle.pair_dist(X, np.repeat(np.arange(nq), kk), nq + idx.ravel()).reshape(nq, kk)
    np.testing.assert_array_equal(dist, pd)


def test_knn_query_edges():
    R = np.arange(12, dtype=np.float32).reshape(6, 2)
    idx, dist = N.knn_query(R, np.zeros((0, 2), np.float32), 3)
    assert idx.shape == (0, 0) and dist.shape == (0, 0)
    idx, dist = N.knn_query(R, R[:2].copy(), 10)  # k > n_ref: all reference rows
    assert idx.shape == (2, 6)
    assert idx[0].tolist() == [0, 1, 2, 3, 4, 5] and dist[0, 0] == 0.0
    with pytest.raises(ValueError):
        N.knn_query(R, np.zeros((3, 3), np.float32), 2)
    with pytest.raises(TypeError):
        N.knn_query(R.astype(np.float64), R, 2)


def _far_clusters(n, p, R, seed):
    """Two tight clusters at +-R along the first axis: every row far from the mean and
    its neighbours close, so the screens' rounding (relative to |x'|^2) exceeds the
    gaps between neighbour distances and the certificate has to recheck."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, p))
    X[:, 0] += np.where(np.arange(n) % 2 == 0, R, -R)
    return X.astype(np.float32)


def _check_graph(oracle, x, idx, dist, k):
    n = x.shape[0]
    assert idx.shape == (n, k) and not (idx == np.arange(n)[:, None]).any()
    ref_i, ref_d2 = oracle.knn_exact_f64(x, k)
    bad, ties = oracle.knn_mismatch_ties(idx, x, np.arange(n), ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(x, np.repeat(np.arange(n), k), idx.ravel()).reshape(n, k)
    np.testing.assert_array_equal(dist, pd)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "fp16"])
def test_knn_certificate_recheck(oracle, prec):
    """Rows 225 sd from the mean: the FP16 and split-BF16 screens cannot prove their
    sets, the FP32 recheck with 128 candidates proves them all, and the sets equal
    the exact FP64 search."""
    n, k = 3000, 15
    x = _far_clusters(n, 20, 225.0, 5)
    ca.set_knn_precision(prec)
    try:
        idx, dist = N.knn_graph(x, k)
        cert = ca.knn_certificate()
    finally:
        ca.set_knn_precision(None)
    assert cert["checked"] == n and cert["uncertified"] == 0, cert
    if prec != "fp64":
        assert cert["rechecked"] > n // 10, cert
    _check_graph(oracle, x, idx, dist, k)


def test_knn_certificate_recheck_wide(oracle):
    """200 columns: the FP16 screen holds them and rechecks what it cannot prove; the
    FP32 GEMM screen of the large-k path certifies its own sets (k = 100)."""
    n, k = 2000, 15
    x = _far_clusters(n, 200, 225.0, 6)
    ca.set_knn_precision("fp16")
    try:
        idx, dist = N.knn_graph(x, k)
        cert = ca.knn_certificate()
        il, dl = N.knn_graph(x, 100)
        cl = ca.knn_certificate()
    finally:
        ca.set_knn_precision(None)
    assert cert["checked"] == n and cert["rechecked"] > 0 and cert["uncertified"] == 0, cert
    _check_graph(oracle, x, idx, dist, k)
    assert cl["checked"] == n and cl["rechecked"] == 0 and cl["uncertified"] <= n // 100, cl
    _check_graph(oracle, x, il, dl, 100)


@pytest.mark.parametrize("d", [30, 50])
@pytest.mark.parametrize("scan", ["mfma", "fp32", "gemm", "valu"])
@pytest.mark.parametrize("prec", ["bf16x3", "fp16"])
def test_knn_certificate_forced_recheck(oracle, prec, scan, d, monkeypatch):
    """A bound scaled by 1e9 proves nothing: every query is rechecked (graph and query
    search) by the threshold scan (after an FP16 screen on its own FP16 operands by
    default; on FP32 MFMA, by rocBLAS or by the VALU kernel),
    whose answers are exact, so none stays unproven."""
    monkeypatch.setenv("CREST_KNN_CERT_SCALE", "1e9")
    monkeypatch.setenv("CREST_KNN_SCAN", scan)
    x = _clustered(2500, d, 11)
    ca.set_knn_precision(prec)
    try:
        idx, dist = N.knn_graph(x, 15)
        cg = ca.knn_certificate()
        R, Q = _clustered(3000, d, 12), (_clustered(700, d, 13) + np.float32(0.5)).astype(np.float32)
        qi, qd = N.knn_query(R, Q, 20)
        cq = ca.knn_certificate()
    finally:
        ca.set_knn_precision(None)
    assert cg == {"checked": 2500, "rechecked": 2500, "uncertified": 0}, cg
    assert cq == {"checked": 700, "rechecked": 700, "uncertified": 0}, cq
    _check_graph(oracle, x, idx, dist, 15)
    ref_i, ref_d2 = oracle.knn_query_exact_f64(R, Q, 20)
    bad, ties = oracle.knn_query_mismatch_ties(qi, R, Q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    X = np.vstack([Q, R])
    pd = oracle.pair_dist(X, np.repeat(np.arange(700), 20), 700 + qi.ravel()).reshape(700, 20)
    np.testing.assert_array_equal(qd, pd)


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_knn_certificate_duplicates(oracle, prec):
    """700 copies of one row: their lists hold only zero distances, the screen cannot
    prove them, the threshold scan finds 699 candidates at distance 0 (more than its
    256 slots) and the GEMM recheck answers them; every set is right up to ties."""
    x = _clustered(2000, 16, 21)
    x[:700] = x[0]
    ca.set_knn_precision(prec)
    try:
        idx, dist = N.knn_graph(x, 15)
        cert = ca.knn_certificate()
    finally:
        ca.set_knn_precision(None)
    assert cert["checked"] == 2000 and cert["rechecked"] >= 700, cert
    assert cert["uncertified"] <= cert["rechecked"], cert
    assert (dist[:700] == 0).all()
    _check_graph(oracle, x, idx, dist, 15)


def test_knn_small_edge_cases():
    idx, dist = N.knn_graph(np.zeros((1, 4), np.float32), 5)
    assert idx.shape == (1, 0)
    idx, dist = N.knn_graph(np.array([[0, 0], [1, 0], [3, 0]], np.float32), 10)
    assert idx.tolist() == [[1, 2], [0, 2], [1, 0]]
    np.testing.assert_array_equal(dist, [[1, 3], [1, 2], [2, 3]])


def test_scaled_pca_matches_dense(counts, oracle):
    """tests/test_crest.py:88-107: implicit scale + Gram PCA == dense scale + SVD."""
    indptr, indices, data, g = counts
    bf = ca.BioFrame.from_csr(indptr, indices, data, g)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    ca.pp.scale(bf, max_value=2.0)
    ca.tl.pca(bf, n_comps=10, use_highly_variable=False)
    D = bf.to_dense().astype(np.float64)
    mu, sd = D.mean(0), D.std(0, ddof=1)
    sd[sd == 0] = 1
    Z = np.clip((D - mu) / sd, -2, 2)
    Z -= Z.mean(0)
    U, S, _ = np.linalg.svd(Z, full_matrices=False)
    ref = (S[:10] ** 2) / (len(Z) - 1)
    np.testing.assert_allclose(bf.uns["pca"]["variance"], ref, rtol=1e-6)
    ref_scores = U[:, :10] * S[:10]
    got = bf.obsm["X_pca"].astype(np.float64)
    cos = np.abs((got * ref_scores).sum(0)) / np.linalg.norm(got, axis=0) / np.linalg.norm(ref_scores, axis=0)
    assert cos.min() > 0.9999


def test_chunking_invariance(counts):
    """tests/test_crest.py:54-61."""
    def pipe(chunk):
        bf = _frame(counts, chunk, drop=False)
        ca.pp.highly_variable_genes(bf, n_top_genes=300)
        ca.pp.scale(bf, max_value=10)
        ca.tl.pca(bf, n_comps=20)
        return bf
    a, b = pipe(None), pipe(1000)
    assert (a.var["highly_variable"] == b.var["highly_variable"]).all()
    np.testing.assert_allclose(np.abs(a.obsm["X_pca"]), np.abs(b.obsm["X_pca"]), rtol=1e-4, atol=1e-3)


def test_hvg_matches_oracle(counts, oracle):
    bf = _frame(counts, drop=False)
    ca.pp.highly_variable_genes(bf, n_top_genes=300)
    st = np.zeros((bf.n_vars, 5))
    for ctx in bf.iter_ctx():
        oracle.gene_stats(*ctx, st)
    n = bf.n_obs
    m = st[:, 2] / n
    v = np.maximum((st[:, 3] - n * m * m) / (n - 1), 0)
    hv, _ = oracle.hvg_seurat_pandas(m, v, n_top_genes=300)
    np.testing.assert_array_equal(bf.var["highly_variable"], hv)


@pytest.mark.parametrize("n,d,k", [(3000, 20, 14), (500, 5, 4)])
def test_connectivities(oracle, n, d, k):
    x = _clustered(n, d, 7 + n)
    idx, dist = N.knn_graph(x, k)
    r, c, w = N.connectivities(idx, dist, k + 1)
    ro, co, wo = oracle.connectivities(idx, dist, k + 1)
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(c, co)
    np.testing.assert_allclose(w, wo, rtol=2e-6, atol=1e-7)


def test_neighbors_api(counts):
    bf = _frame(counts, drop=False)
    ca.pp.highly_variable_genes(bf, n_top_genes=200)
    ca.pp.scale(bf)
    ca.tl.pca(bf, n_comps=15)
    ca.pp.neighbors(bf, n_neighbors=10)
    nb = bf.uns["neighbors"]
    assert nb["indices"].shape == (bf.n_obs, 9) and len(nb["rows"]) > 0
    C = ca.pp.connectivities_matrix(bf)
    assert (abs(C - C.T) > 0).nnz == 0


def test_parquet_roundtrip(counts, tmp_path, oracle):
    """write_parquet of a filtered frame, read back: the same raw counts (QC bit-exact)
    and the same transformed statistics, one chunk per part file."""
    bf = _frame(counts, 4000)
    bf.write_parquet(tmp_path / "ds", nnz_per_part=7000)
    back = ca.read_parquet(tmp_path / "ds")
    assert len(back.store.parts) > 1 and back.n_obs == bf.n_obs and back.n_vars == bf.n_vars
    D0 = bf.to_dense(transform=False)
    D1 = back.to_dense(transform=False)
    np.testing.assert_array_equal(D0, D1)
    ca.pp.normalize_total(back, 1e4)
    ca.pp.log1p(back)
    np.testing.assert_allclose(ca.pp.gene_stats(back), ca.pp.gene_stats(bf), rtol=1e-12, atol=1e-12)
    o0, v0 = ca.pp.calculate_qc_metrics(bf, inplace=False)
    o1, v1 = ca.pp.calculate_qc_metrics(back, inplace=False)
    np.testing.assert_array_equal(o0["total_counts"], o1["total_counts"])
    np.testing.assert_array_equal(v0["n_cells_by_counts"], v1["n_cells_by_counts"])


def _gram_with_spectrum(lam, n, seed):
    """Gram and colsum (zero) whose covariance (G - c c^T / n) / (n - 1) has eigenvalues lam."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(len(lam), len(lam))))
    C = (Q * lam) @ Q.T
    C = (C + C.T) / 2
    return C * (n - 1), np.zeros(len(lam)), C


def _subspace_angle(A, B):
    s = np.linalg.svd(A.T @ B, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1))) if s.min() < 1 - 1e-14 else float(
        np.linalg.norm(B - A @ (A.T @ B), 2))


@pytest.mark.parametrize("case", ["cluster", "rank_deficient", "dense_distinct"])
def test_pca_from_gram_top_p_eigensolver(case, monkeypatch):
    """The top-p eigensolver (tridiagonalization, bisection, inverse iteration with
    in-cluster orthogonalisation, back-transform) on spectra with an exactly repeated
    cluster, with zero eigenvalues inside the requested p, and fully distinct: the
    eigenvalues of numpy's eigh, orthonormal loadings, the determined subspaces, and
    the same results as rocSOLVER's full dsyevd."""
    d, n, k = 300, 5000, 40
    rng = np.random.default_rng(3)
    if case == "cluster":
        lam = np.concatenate([np.linspace(10, 1, 10), np.full(30, 0.5), np.sort(rng.uniform(0, 0.3, d - 40))[::-1]])
    elif case == "rank_deficient":
        lam = np.concatenate([np.sort(rng.uniform(1, 5, 25))[::-1], np.zeros(d - 25)])
    else:
        lam = np.sort(rng.uniform(0.01, 20, d))[::-1]
    g, c, C = _gram_with_spectrum(lam, n, 7)
    monkeypatch.setenv("CREST_EIGENSOLVER", "topp")
    var, load, tot = N.pca_from_gram(g, c, n, k)
    ev, V = np.linalg.eigh(C)
    ev, V = ev[::-1], V[:, ::-1]
    np.testing.assert_allclose(var, np.maximum(ev[:k], 0), atol=1e-12 * ev[0])
    assert abs(tot - np.trace(C)) <= 1e-12 * abs(np.trace(C))
    assert np.abs(load.T @ load - np.eye(k)).max() < 1e-12
    # residual of every computed pair
    assert np.abs(C @ load - load * var).max() < 1e-11 * ev[0]
    if case == "cluster":
        assert _subspace_angle(V[:, :10], load[:, :10]) < 1e-10
        assert _subspace_angle(V[:, :40], load[:, :40]) < 1e-10
    elif case == "dense_distinct":
        assert _subspace_angle(V[:, :k], load) < 1e-9
    else:
        assert _subspace_angle(V[:, :25], load[:, :25]) < 1e-10
    # sign rule: the first largest-|loading| entry of every component is positive
    idx = np.argmax(np.abs(load), axis=0)
    assert (load[idx, np.arange(k)] > 0).all()
    monkeypatch.setenv("CREST_EIGENSOLVER", "dsyevd")
    var2, load2, tot2 = N.pca_from_gram(g, c, n, k)
    np.testing.assert_allclose(var, var2, atol=1e-12 * ev[0])
    assert tot == tot2
    if case == "dense_distinct":
        np.testing.assert_allclose(load, load2, atol=1e-9)


@pytest.mark.parametrize("case,d,k", [
    ("flat_tail", 1200, 50), ("dominant", 1200, 50), ("cluster_at_cut", 1000, 40), ("dense_distinct", 600, 64),
    ("rank_deficient", 800, 50), ("small_forced", 300, 20)])
def test_pca_from_gram_filtered_subspace(case, d, k, monkeypatch, capfd):
    """The Chebyshev-filtered subspace iteration (the default eigensolver from order
    512) on spectra shaped like a scaled expression matrix's (one dominant value and a
    flat tail with gaps of 1e-3 around the cut), with a dominant block, with an exactly
    repeated cluster across the cut, fully distinct with the largest block, and rank
    deficient below p (the iteration gives up and the tridiagonal path answers): the
    eigenvalues of numpy's eigh, orthonormal loadings, residuals, the determined
    subspaces and the sign rule, and the tridiagonal path's results."""
    rng = np.random.default_rng(11)
    if case == "flat_tail":
        lam = np.concatenate([[17.4, 4.7, 4.5], np.sort(rng.uniform(1.3, 3.0, 40))[::-1],
                              np.linspace(1.28, 1.20, 60), np.sort(rng.uniform(0.0, 1.2, d - 103))[::-1]])
    elif case == "dominant":
        lam = np.concatenate([np.geomspace(1e4, 50, 10), np.sort(rng.uniform(1, 40, 60))[::-1],
                              np.sort(rng.uniform(0, 1, d - 70))[::-1]])
    elif case == "cluster_at_cut":
        lam = np.concatenate([np.linspace(20, 5, 35), np.full(12, 3.0), np.sort(rng.uniform(0, 2.5, d - 47))[::-1]])
    elif case == "dense_distinct":
        lam = np.sort(rng.uniform(0.01, 20, d))[::-1]
    elif case == "rank_deficient":
        lam = np.concatenate([np.sort(rng.uniform(1, 5, 30))[::-1], np.zeros(d - 30)])
    else:
        lam = np.sort(rng.uniform(0.01, 20, d))[::-1]
    g, c, C = _gram_with_spectrum(lam, 5000, 5)
    monkeypatch.setenv("CREST_CHEFSI_TRACE", "1")
    if case == "small_forced":
        monkeypatch.setenv("CREST_EIGENSOLVER", "chefsi")
    else:
        monkeypatch.delenv("CREST_EIGENSOLVER", raising=False)
    print(f"[chefsi] {case}: d {d}, p {k}", flush=True)
    var, load, tot = N.pca_from_gram(g, c, 5000, k)
    err = capfd.readouterr().err
    if case == "rank_deficient":
        assert "[chefsi] fallback" in err
    else:
        assert "[chefsi] converged" in err, err[-2000:]
    ev, V = np.linalg.eigh(C)
    ev, V = ev[::-1], V[:, ::-1]
    np.testing.assert_allclose(var, np.maximum(ev[:k], 0), atol=1e-12 * ev[0], rtol=1e-10)
    assert abs(tot - np.trace(C)) <= 1e-12 * abs(np.trace(C))
    assert np.abs(load.T @ load - np.eye(k)).max() < 1e-12
    assert np.abs(C @ load - load * var).max() < 1e-10 * ev[0]
    # every leading subspace that a relative gap of 1e-4 separates
    for r in range(1, k + 1):
        gap = ev[r - 1] - ev[r]
        if gap > 1e-4 * ev[0]:
            assert _subspace_angle(V[:, :r], load[:, :r]) < 1e-9 * ev[0] / gap, r
    idx = np.argmax(np.abs(load), axis=0)
    assert (load[idx, np.arange(k)] > 0).all()
    monkeypatch.setenv("CREST_EIGENSOLVER", "topp")
    var2, load2, _ = N.pca_from_gram(g, c, 5000, k)
    np.testing.assert_allclose(var, var2, atol=1e-12 * ev[0], rtol=1e-10)
    iso = np.minimum(np.abs(np.diff(np.r_[np.inf, ev[:k]])), np.abs(np.diff(ev[:k + 1]))) > 1e-2 * ev[0]
    for j in np.nonzero(iso)[0]:  # isolated components: the same vector, the same sign
        assert np.abs(load[:, j] - load2[:, j]).max() < 1e-8


def test_log1pf_device_matches_libm(oracle):
    """The device transform's log1pf (log1pf_glibc.h) against the oracle's, which calls
    the C library's log1pf (the reference's f32::ln_1p, kernels.rs:81): every 97th
    float of [-1, +inf] and every float within 2^16 of each branch threshold, bit for
    bit (no normalisation, so the transform is log1pf of the raw value)."""
    pos = np.arange(0, 0x7F800001, 97, dtype=np.int64)
    neg = np.arange(0x80000000, 0xBF800001, 97, dtype=np.int64)
    edges = [0x3ED413D7, 0x3504F7 | 0x3F800000, 0xBE95F61F, 0x31000000, 0x24800000, 0x5A000000, 0xB1000000,
             0x3F800000, 0xBF800000 - 65536]
    near = np.concatenate([np.arange(e - 65536, e + 65536, dtype=np.int64) for e in edges])
    bits = np.unique(np.concatenate([pos, neg, near, [0x7F800000, 0xBF800000]]))
    x = bits.astype(np.uint32).view(np.float32)
    x = x[(x >= -1.0) | ~np.isfinite(x)]
    x = x[~np.isnan(x)]
    per = 64
    n = len(x) // per * per
    x = np.ascontiguousarray(x[:n])
    cells = n // per
    genes = np.tile(np.arange(per, dtype=np.uint32), cells)
    ip = np.arange(0, n + 1, per, dtype=np.int64)
    ctx = (genes, x, np.arange(cells, dtype=np.int64), np.arange(per, dtype=np.int32), 0.0, True, None, ip, 0)
    print(f"[log1pf] {n} floats through crest_materialize", flush=True)
    got = N.materialize(*ctx)[2]
    ref = oracle.materialize(*ctx)[2]
    assert len(got) == len(ref) == n
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("dense_rows", [False, True])
def test_sparse_fp64_gram(oracle, dense_rows, monkeypatch):
    """The FP64 Gram by sparse pair products (the default FP64 form) against the oracle
    (kernels.rs:221-309) and against the dense FP64 MFMA form (CREST_GRAM_SPARSE=0), to
    FP64 rounding; with rows whose every gene is non-zero, a group's bucket exceeds the
    staged capacity and the unstaged branch runs."""
    ip, ix, dv, _ = synth.make_counts(5000, 700, 6, density=0.08, seed=31)
    if dense_rows:  # cells 100..169 express every gene
        rows = [np.arange(ip[r], ip[r + 1]) for r in range(len(ip) - 1)]
        new_ix, new_dv, new_ip = [], [], [0]
        rng = np.random.default_rng(5)
        for r, sl in enumerate(rows):
            if 100 <= r < 170:
                g = np.arange(700, dtype=np.uint32)
                v = rng.integers(1, 20, 700).astype(np.float32)
            else:
                g, v = ix[sl], dv[sl]
            new_ix.append(g)
            new_dv.append(v)
            new_ip.append(new_ip[-1] + len(g))
        ip, ix, dv = np.asarray(new_ip, np.int64), np.concatenate(new_ix), np.concatenate(new_dv)
    bf = ca.BioFrame.from_csr(ip, ix, dv, 700)
    bf.chunk_nnz = len(dv) // 3
    rng = np.random.default_rng(0)
    bf = bf.filter_cells(rng.random(bf.n_obs) > 0.05).filter_genes(rng.random(bf.n_vars) > 0.1)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    gmap, mean, sd, d = _gram_inputs(bf, oracle, d_top=600)
    got = [np.zeros((d, d)), np.zeros(d)]
    ref = [np.zeros((d, d)), np.zeros(d)]
    dense = [np.zeros((d, d)), np.zeros(d)]
    ca.set_precision("fp64")
    for ctx in bf.iter_ctx():
        N.gram_accumulate(*ctx, gmap, got[0], got[1], True, mean, sd, 10.0)
        oracle.gram_accumulate(*ctx, gmap, ref[0], ref[1], True, mean, sd, 10.0)
    monkeypatch.setenv("CREST_GRAM_SPARSE", "0")
    for ctx in bf.iter_ctx():
        N.gram_accumulate(*ctx, gmap, dense[0], dense[1], True, mean, sd, 10.0)
    iu = np.triu_indices(d)
    scale = np.abs(ref[0][iu]).max()
    assert np.abs(got[0][iu] - ref[0][iu]).max() <= 1e-12 * scale
    assert np.abs(got[0][iu] - dense[0][iu]).max() <= 1e-12 * scale
    assert np.all(np.tril(got[0], -1) == 0)
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-12, atol=1e-9)
