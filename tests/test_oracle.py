"""Pins of the oracle (CPU): the reference's own known answers and tests restated,
closed forms, dense FP64 references, brute force, invariances, and checks that
plausible mistakes would fail these pins."""
import numpy as np
import pytest
import scipy.sparse as sp


def make_counts(n_cells=400, n_genes=120, n_clusters=4, seed=0):
    """tests/test_crest.py:13-27 generator (NB counts with cluster programs)."""
    rng = np.random.default_rng(seed)
    base = rng.gamma(0.4, 1.0, n_genes)
    labels = rng.integers(0, n_clusters, n_cells)
    prog = np.ones((n_clusters, n_genes))
    for c in range(n_clusters):
        up = rng.choice(n_genes, 12, replace=False)
        prog[c, up] *= rng.uniform(3, 10, 12)
    lib = rng.lognormal(7.5, 0.4, n_cells)
    mu = prog[labels] * base[None, :]
    mu = mu / mu.sum(1, keepdims=True) * lib[:, None]
    counts = rng.negative_binomial(2, 2 / (2 + mu)).astype(np.float32)
    return sp.csr_matrix(counts)


@pytest.fixture(scope="module")
def X():
    return make_counts()


def ctx_of(X, target=1e4, log=True, cell_map=None, gene_map=None, rows=None):
    X = X.tocsr()
    X.sort_indices()
    n, g = X.shape
    lo, hi = (0, n) if rows is None else rows
    ip = X.indptr.astype(np.int64)
    cm = np.arange(n, dtype=np.int64) if cell_map is None else cell_map
    gm = np.arange(g, dtype=np.int32) if gene_map is None else gene_map
    a, b = ip[lo], ip[hi]
    return (X.indices[a:b].astype(np.uint32), X.data[a:b].astype(np.float32), cm, gm, target, log, None,
            ip[lo:hi + 1].copy(), lo)


def test_materialize_hand_values(oracle):
    """kernels.rs:564-578."""
    cells = np.array([0, 0, 1, 2, 2], np.uint32)
    genes = np.array([0, 1, 0, 1, 2], np.uint32)
    vals = np.array([2, 6, 5, 1, 3], np.float32)
    cm = np.array([0, -1, 1], np.int64)
    gm = np.array([0, 1, -1], np.int32)
    r, g, x = oracle.materialize(genes, vals, cm, gm, 100.0, False, cells, None, 0)
    assert r.tolist() == [0, 0, 1] and g.tolist() == [0, 1, 1]
    assert np.allclose(x, [25, 75, 100], atol=1e-5)
    ip = np.array([0, 2, 3, 5], np.int64)
    r2, g2, x2 = oracle.materialize(genes, vals, cm, gm, 100.0, False, None, ip, 0)
    assert r2.tolist() == r.tolist() and g2.tolist() == g.tolist() and x2.tolist() == x.tolist()


def _libm_log1pf(y: np.ndarray) -> np.ndarray:
    """The C library's log1pf on each float32 of y, called one by one through ctypes
    (the reference's f32::ln_1p, kernels.rs:81), independent of the oracle's C code."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.log1pf.restype = ctypes.c_float
    libm.log1pf.argtypes = [ctypes.c_float]
    return np.array([libm.log1pf(float(v)) for v in np.asarray(y, np.float32)], np.float32)


def test_lazy_normalisation_closed_form(oracle, X):
    """tests/test_crest.py:77-85, and bit-exact against the reference's float32
    rounding chain y = (f32)(x * (1e4/total)), v = log1pf(y)."""
    r, g, x = oracle.materialize(*ctx_of(X))
    D = X.toarray().astype(np.float64)
    ref = np.log1p(D / D.sum(1, keepdims=True) * 1e4)
    np.testing.assert_allclose(x, ref[r, g], rtol=1e-5, atol=1e-5)
    tot = D.sum(1)
    y = (D[r, g] * (1e4 / tot[r])).astype(np.float32)
    np.testing.assert_array_equal(x, _libm_log1pf(y))


def test_log1p_modes_differ_by_at_most_one_ulp(oracle, X):
    """Mode 0 (correctly rounded float of the double log1p) against the default mode 1
    (log1pf): at most one ulp apart, and mode 0 equals the correctly rounded chain."""
    ctx = ctx_of(X)
    _, _, b = oracle.materialize(*ctx)
    oracle.set_log1p_mode(0)
    try:
        r, g, a = oracle.materialize(*ctx)
    finally:
        oracle.set_log1p_mode(1)
    ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    D = X.toarray().astype(np.float64)
    y = (D[r, g] * (1e4 / D.sum(1)[r])).astype(np.float32)
    np.testing.assert_array_equal(a, np.log1p(y.astype(np.float64)).astype(np.float32))


def test_filters_and_maps(oracle, X):
    n, g = X.shape
    rng = np.random.default_rng(2)
    keep_c = rng.random(n) > 0.2
    keep_g = rng.random(g) > 0.3
    cm = np.full(n, -1, np.int64)
    cm[keep_c] = np.arange(keep_c.sum())
    gm = np.full(g, -1, np.int32)
    gm[keep_g] = np.arange(keep_g.sum())
    r, v, x = oracle.materialize(*ctx_of(X, cell_map=cm, gene_map=gm))
    D = X.toarray().astype(np.float64)[keep_c][:, keep_g]
    ref = np.log1p(D / D.sum(1, keepdims=True) * 1e4)
    np.testing.assert_allclose(x, ref[r, v], rtol=1e-5, atol=1e-5)
    assert len(x) == np.count_nonzero(D)


def lognorm(X):
    """The reference's rounding chain: y = (f32)(x * (1e4 / total)), v = log1pf(y) (dense)."""
    D = X.toarray().astype(np.float64)
    y = (D * (1e4 / D.sum(1, keepdims=True))).astype(np.float32)
    out = np.zeros(y.shape, np.float64)
    nz = y != 0
    out[nz] = _libm_log1pf(y[nz]).astype(np.float64)
    return out


def test_gene_stats_dense(oracle, X):
    st = np.zeros((X.shape[1], 5))
    oracle.gene_stats(*ctx_of(X), st)
    L = lognorm(X)
    E = np.expm1(L)
    ref = np.stack([L.sum(0), (L * L).sum(0), E.sum(0), (E * E).sum(0), (L != 0).sum(0)], 1)
    np.testing.assert_allclose(st, ref, rtol=1e-12, atol=1e-12)


def _scaled(X, mv):
    L = lognorm(X)
    n = len(L)
    mean = L.mean(0)
    sd = np.sqrt(np.maximum(((L * L).sum(0) - n * mean * mean) / (n - 1), 0))
    sd[sd == 0] = 1
    return L, mean, sd


def test_gram_matches_dense(oracle, X):
    """G = S^T S with S = clip((x-mu)/sd) - clip(-mu/sd) on non-zeros (kernels.rs:197-219)."""
    mv = 2.0
    L, mean, sd = _scaled(X, mv)
    d = X.shape[1]
    gmap = np.arange(d, dtype=np.int32)
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    oracle.gram_accumulate(*ctx_of(X), gmap, gram, colsum, True, mean, sd, mv)
    Z = np.clip((L - mean) / sd, -mv, mv)
    b = np.clip(-mean / sd, -mv, mv)
    S = np.where(L != 0, Z - b, 0.0)
    ref = S.T @ S
    iu = np.triu_indices(d)
    np.testing.assert_allclose(gram[iu], ref[iu], rtol=1e-11, atol=1e-9)
    assert np.all(np.tril(gram, -1) == 0)
    np.testing.assert_allclose(colsum, S.sum(0), rtol=1e-12, atol=1e-9)
    # a plausible mistake -- dropping the clipped-zero offset b -- is caught
    wrong = np.where(L != 0, Z, 0.0)
    assert not np.allclose(gram[iu], (wrong.T @ wrong)[iu], rtol=1e-6)


def test_gram_chunk_and_thread_invariance(oracle, X):
    L, mean, sd = _scaled(X, 10.0)
    d = X.shape[1]
    gmap = np.arange(d, dtype=np.int32)
    g1, c1 = np.zeros((d, d)), np.zeros(d)
    oracle.gram_accumulate(*ctx_of(X), gmap, g1, c1, True, mean, sd, 10.0, threads=1)
    g4, c4 = np.zeros((d, d)), np.zeros(d)
    oracle.gram_accumulate(*ctx_of(X), gmap, g4, c4, True, mean, sd, 10.0, threads=4)
    np.testing.assert_array_equal(g1, g4)
    gc, cc = np.zeros((d, d)), np.zeros(d)
    n = X.shape[0]
    for lo, hi in [(0, 130), (130, 131), (131, n)]:
        oracle.gram_accumulate(*ctx_of(X, rows=(lo, hi)), gmap, gc, cc, True, mean, sd, 10.0)
    np.testing.assert_allclose(gc, g1, rtol=1e-12, atol=1e-10)


def test_scaled_pca_matches_dense_svd(oracle, X):
    """tests/test_crest.py:88-107 and kernels.rs:580-633 (eigenvalues of the scaled,
    clipped, centred dense matrix)."""
    mv = 2.0
    L, mean, sd = _scaled(X, mv)
    n, d = L.shape
    gmap = np.arange(d, dtype=np.int32)
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    oracle.gram_accumulate(*ctx_of(X), gmap, gram, colsum, True, mean, sd, mv)
    var, load, total = oracle.pca_from_gram(gram, colsum, n, 10)
    Z = np.clip((L - mean) / sd, -mv, mv)
    Zc = Z - Z.mean(0)
    U, S, Vt = np.linalg.svd(Zc, full_matrices=False)
    np.testing.assert_allclose(var, S[:10] ** 2 / (n - 1), rtol=1e-9)
    np.testing.assert_allclose(total, (Zc * Zc).sum() / (n - 1), rtol=1e-12)
    assert oracle.principal_angle(load[:, :5], Vt[:5].T) < 1e-8
    # projection: X = -shift + Σ s V == Zc V
    shift = (colsum / n) @ load
    Xp = np.empty((n, 10), np.float32)
    Xp[:] = -shift.astype(np.float32)
    oracle.project(*ctx_of(X), gmap, load, Xp, True, mean, sd, mv)
    np.testing.assert_allclose(Xp, Zc @ load, rtol=1e-4, atol=1e-4)


def test_sign_rule(oracle):
    """kernels.rs:334-345: the first largest-|loading| entry of each component is
    positive."""
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    Q[:, 0] = [0.1, -0.7, 0.7, 0.1, 0.0, 0.0]
    Q, _ = np.linalg.qr(Q)
    lam = np.array([9.0, 5.0, 3.0, 2.0, 1.0, 0.5])
    cov = (Q * lam) @ Q.T
    n = 11
    gram = np.triu(cov * (n - 1))
    var, load, total = oracle.pca_from_gram(gram, np.zeros(6), n, 6)
    np.testing.assert_allclose(var, lam, rtol=1e-12)
    for c in range(6):
        i = int(np.argmax(np.abs(load[:, c])))
        assert load[i, c] > 0
        np.testing.assert_allclose(np.abs(load[:, c]), np.abs(Q[:, c]), atol=1e-10)


def test_pca_from_gram_errors(oracle):
    with pytest.raises(RuntimeError):
        oracle.pca_from_gram(np.zeros((2, 2)), np.zeros(2), 1, 1)


def _clustered(n, d, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-10, 10, (20, d))
    return (centers[np.arange(n) % 20] + rng.uniform(-1.5, 1.5, (n, d))).astype(np.float32)


def test_knn_brute_matches_naive(oracle):
    """knn.rs:399-414 and tests/test_crest.py:504-511."""
    x = _clustered(600, 7, 1)
    idx, dist = oracle.knn(x, 10)
    D = ((x[:, None, :].astype(np.float64) - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    order = np.argsort(D, axis=1, kind="stable")[:, :10]
    assert not (idx == np.arange(600)[:, None]).any()
    np.testing.assert_allclose(dist[:, 9], np.sqrt(np.sort(D, 1)[:, 9]), rtol=1e-4)
    same = (np.sort(idx, 1) == np.sort(order, 1)).all(1).mean()
    assert same > 0.99
    rng = np.random.default_rng(0)
    y = rng.normal(size=(800, 10)).astype(np.float32)
    idx, dist = oracle.knn(y, 5)
    D = ((y[:, None, :].astype(np.float64) - y[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    np.testing.assert_array_equal(idx, np.argsort(D, axis=1)[:, :5])
    np.testing.assert_allclose(dist, np.sqrt(np.sort(D, axis=1)[:, :5]), rtol=1e-4)


def test_knn_exact_f64_helper(oracle):
    y = np.random.default_rng(3).normal(size=(500, 6)).astype(np.float32)
    i, d2 = oracle.knn_exact_f64(y, 4, np.arange(0, 500, 7))
    y = y.astype(np.float64)
    D = ((y[::7, None, :] - y[None]) ** 2).sum(-1)
    D[np.arange(len(D)), np.arange(0, 500, 7)] = np.inf
    np.testing.assert_array_equal(i, np.argsort(D, 1)[:, :4])
    np.testing.assert_allclose(d2, np.sort(D, 1)[:, :5], rtol=1e-12)


def test_principal_angle(oracle):
    t = 1e-7
    a = np.array([[1.0, 0], [0, 1], [0, 0]])
    b = np.array([[1.0, 0], [0, np.cos(t)], [0, np.sin(t)]])
    assert abs(oracle.principal_angle(a, b) - t) < 1e-12
    assert oracle.principal_angle(a, a[:, ::-1]) < 1e-15


def test_cell_ranger_columns_match_pandas_oracle(oracle):
    """The cell_ranger flavour (host numpy from the device statistics) against a pandas
    restatement: percentile bins (pandas.cut), groupby median and MAD."""
    from crest_amd.pp import _cell_ranger_columns
    Xb = make_counts(1500, 600, 5, seed=4)
    st = np.zeros((600, 5))
    oracle.gene_stats(*ctx_of(Xb), st)
    n = Xb.shape[0]
    for n_top in (200, None):
        cols = _cell_ranger_columns(st, n, n_top, 0.0125, 3.0, 0.5, np.inf)
        hv, dn = oracle.hvg_cell_ranger_pandas(st[:, 0] / n, st, n, n_top_genes=n_top)
        np.testing.assert_allclose(cols["dispersions_norm"], dn, rtol=1e-12, equal_nan=True)
        np.testing.assert_array_equal(cols["highly_variable"], hv)


def test_connectivities_pins(oracle):
    """umap/graph.rs:127-160: the 4-point toy graph is symmetric and links the two
    pairs; each row's membership strengths sum to log2(n_neighbors)."""
    x = np.array([[0, 0], [0.1, 0.1], [1, 1], [1.1, 1.1]], np.float32)
    idx, dist = oracle.knn(x, 1)
    r, c, w = oracle.connectivities(idx, dist, 2)
    pairs = set(zip(r.tolist(), c.tolist()))
    assert (0, 1) in pairs and (2, 3) in pairs and np.all(r < c)
    # sigma hits the target: rebuild the per-row strengths from the edge list
    rng = np.random.default_rng(0)
    y = rng.normal(size=(300, 5)).astype(np.float32)
    idx, dist = oracle.knn(y, 9)
    r, c, w = oracle.connectivities(idx, dist, 10)
    assert np.all(r < c) and np.all(np.diff(r.astype(np.int64) * 300 + c) > 0)
    assert np.all((w > 0) & (w <= 1))
    # directed strengths of row 0 from its own (rho, sigma): Σ = log2(10) within 1e-4
    d = dist[0].astype(np.float64)
    rho = d[d > 0][0]
    lo, hi = 1e-6, 100.0
    for _ in range(200):
        mid = (lo + hi) / 2
        s = np.where(d - rho > 0, np.exp(-(d - rho) / mid), 1.0).sum()
        lo, hi = (lo, mid) if s > np.log2(10) else (mid, hi)
    assert abs(np.where(d - rho > 0, np.exp(-(d - rho) / lo), 1.0).sum() - np.log2(10)) < 1e-4
    r0, c0, w0 = oracle.connectivities(np.zeros((0, 3), np.uint32), np.zeros((0, 3), np.float32), 4)
    assert len(r0) == 0


def test_knn_query_oracle_pins(oracle):
    """tests/test_crest.py:504-511 restated on the oracle's exact query search: the
    indices equal numpy's argsort of the full distance matrix; an off-by-one in the
    partition or a transposed operand would change them."""
    rng = np.random.default_rng(0)
    R = rng.normal(size=(3000, 10)).astype(np.float32)
    Q = rng.normal(size=(200, 10)).astype(np.float32)
    idx, d2 = oracle.knn_query_exact_f64(R, Q, 5)
    D = ((Q[:, None, :].astype(np.float64) - R[None, :, :]) ** 2).sum(-1)
    np.testing.assert_array_equal(idx, np.argsort(D, axis=1, kind="stable")[:, :5])
    np.testing.assert_allclose(d2, np.sort(D, axis=1)[:, :6], rtol=1e-12)
    # k == n_ref: every reference row, the (k+1)-th slot is +inf
    idx, d2 = oracle.knn_query_exact_f64(R[:4], Q[:3], 4)
    assert sorted(idx[0].tolist()) == [0, 1, 2, 3] and np.isinf(d2[:, 4]).all()


def _qc_ctx(indptr, indices, data, n_cells, n_genes, cmap=None, gmap=None, lo=0, hi=None):
    hi = n_cells if hi is None else hi
    a, b = int(indptr[lo]), int(indptr[hi])
    cmap = np.arange(n_cells, dtype=np.int64) if cmap is None else cmap
    gmap = np.arange(n_genes, dtype=np.int32) if gmap is None else gmap
    return (np.ascontiguousarray(indices[a:b], np.uint32), np.ascontiguousarray(data[a:b], np.float32), cmap, gmap,
            0.0, False, None, np.ascontiguousarray(indptr[lo:hi + 1], np.int64), lo)


def test_qc_oracle_pins(oracle):
    """kernels.rs:162-195 on a hand-checked toy: explicit zeros are skipped, a filtered
    gene and a filtered cell contribute nothing, flagged genes feed cell_flag."""
    # cells x genes (gene 0 is "MT-"); cell 1 is filtered, gene 3 is filtered
    indptr = np.array([0, 3, 5, 9], np.int64)
    indices = np.array([0, 1, 3, 0, 2, 0, 1, 2, 3], np.uint32)
    data = np.array([2, 0, 7, 5, 1, 1, 4, 3, 9], np.float32)
    cmap = np.array([0, -1, 1], np.int64)
    gmap = np.array([0, 1, 2, -1], np.int32)
    flag = np.array([1, 0, 0], np.uint8)
    ct, cn, cf = np.zeros(2), np.zeros(2, np.uint32), np.zeros(2)
    gn, gt = np.zeros(3, np.uint32), np.zeros(3)
    oracle.qc(*_qc_ctx(indptr, indices, data, 3, 4, cmap, gmap), flag, ct, cn, cf, gn, gt)
    assert ct.tolist() == [2.0, 8.0] and cn.tolist() == [1, 3] and cf.tolist() == [2.0, 1.0]
    assert gn.tolist() == [2, 1, 1] and gt.tolist() == [3.0, 4.0, 3.0]
    # random integer counts: equals the dense numpy reductions, and chunking adds up
    rng = np.random.default_rng(5)
    D = rng.poisson(0.7, size=(300, 40)).astype(np.float32)
    D[rng.random(D.shape) < 0.3] = 0
    nz = D != 0
    indptr = np.concatenate([[0], np.cumsum(nz.sum(1))]).astype(np.int64)
    indices = np.nonzero(nz)[1].astype(np.uint32)
    data = D[nz]
    flag = (np.arange(40) % 7 == 0).astype(np.uint8)
    out = [np.zeros(300), np.zeros(300, np.uint32), np.zeros(300), np.zeros(40, np.uint32), np.zeros(40)]
    for lo in range(0, 300, 70):
        oracle.qc(*_qc_ctx(indptr, indices, data, 300, 40, lo=lo, hi=min(300, lo + 70)), flag, *out)
    np.testing.assert_array_equal(out[0], D.sum(1))
    np.testing.assert_array_equal(out[1], nz.sum(1))
    np.testing.assert_array_equal(out[2], D[:, flag.astype(bool)].sum(1))
    np.testing.assert_array_equal(out[3], nz.sum(0))
    np.testing.assert_array_equal(out[4], D.sum(0))
