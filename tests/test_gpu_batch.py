"""The batch of independent subsets (crest_pipeline_run_batched) against one
crest_pipeline_run per subset and against the oracle: variance, loadings, scores and
kNN of every subset, with subsets cut across tiles, split into several Gram row
ranges, eigensolved in small batches, and replayed from captured HIP graphs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crest_amd import synth  # noqa: E402

G = 400
# a subset of exactly max(n_comps, k) + 1 cells, ones below and above one 512-row
# query block, and ones spanning several blocks and Gram row ranges
SIZES = [31, 200, 513, 700, 1300, 2600, 4100]


@pytest.fixture(scope="module")
def batch():
    indptr, indices, data, off = synth.make_subsets(SIZES, G, 10, density=0.07, seed=21)
    return indptr, indices, data, off


def _run(batch, **kw):
    from crest_amd.pipeline import BatchPipeline, DeviceCSR
    indptr, indices, data, off = batch
    csr = DeviceCSR.upload(indptr, indices, data, G)
    p = BatchPipeline(csr, off, np.arange(G, dtype=np.int32), n_comps=30, k=15, **kw)
    out = p.run()
    idx, dist = p.knn()
    return p, out, p.x_pca(), idx, dist


def _single(batch, i, prec_gram, prec_knn):
    from crest_amd.pipeline import DeviceCSR, Pipeline
    indptr, indices, data, off = batch
    a, b = int(off[i]), int(off[i + 1])
    ip = indptr[a:b + 1] - indptr[a]
    ix = indices[indptr[a]:indptr[b]]
    dv = data[indptr[a]:indptr[b]]
    p = Pipeline(DeviceCSR.upload(ip, ix, dv, G), np.arange(G, dtype=np.int32), n_comps=30, k=15,
                 prec_gram=prec_gram, prec_knn=prec_knn, chunk_rows=2048)
    out = p.run()
    idx, dist = p.knn()
    return out, p.x_pca(), idx, dist, (ip, ix, dv)


def _angle(oracle, A, B, r):
    return oracle.principal_angle(A[:, :r], B[:, :r])


def _check_knn(oracle, X, idx, n):
    assert idx.max() < n
    assert not (idx == np.arange(n)[:, None]).any()
    q = np.arange(n)
    ref_i, ref_d2 = oracle.knn_exact_f64(X, 15, q)
    bad, ties = oracle.knn_mismatch_ties(idx, X, q, ref_i, ref_d2)
    assert bad == ties, (n, bad, ties)


# Split precisions accumulate 16 K-steps in FP32 before adding into FP64; the batch's
# row ranges move those flush points, so its sums differ from the single run's at the
# FP32 accumulation level (the same bound the single pipeline meets against the oracle).
@pytest.mark.parametrize("prec_gram,prec_knn,var_rtol,tot_rtol", [("fp64", "fp32", 1e-9, 1e-9),
                                                                  ("fp16", "fp16", 1e-6, 1e-7),
                                                                  ("bf16x3", "bf16x3", 1e-5, 1e-6)])
def test_batch_equals_single_runs(batch, oracle, prec_gram, prec_knn, var_rtol, tot_rtol):
    p, out, X, idx, dist = _run(batch, prec_gram=prec_gram, prec_knn=prec_knn)
    off = batch[3]
    for i in range(len(SIZES)):
        a, b = int(off[i]), int(off[i + 1])
        n = b - a
        o1, X1, i1, d1, csr1 = _single(batch, i, prec_gram, prec_knn)
        np.testing.assert_allclose(out.variance[i], o1.variance, rtol=var_rtol, err_msg=f"subset {i}")
        assert abs(out.total_variance[i] - o1.total_variance) <= tot_rtol * o1.total_variance
        # kNN of the batch is exact on the batch's own scores
        _check_knn(oracle, X[a:b], idx[a:b].astype(np.int64), n)
        # the oracle's PCA of the subset
        ip, ix, dv = csr1
        cmap, gmap = np.arange(n, dtype=np.int64), np.arange(G, dtype=np.int32)
        ctx = (ix, dv, cmap, gmap, 1e4, True, None, ip, 0)
        st = np.zeros((G, 5))
        oracle.gene_stats(*ctx, st)
        mean = st[:, 0] / n
        sd = np.sqrt(np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0))
        sd[sd == 0] = 1
        gram, colsum = np.zeros((G, G)), np.zeros(G)
        oracle.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=4)
        var_o, load_o, tot_o = oracle.pca_from_gram(gram, colsum, n, 30)
        # the 10 leading components: the subsets have 10 clusters, so the signal
        # subspace is well separated from the noise.  Subsets of at most 256 cells (fewer
        # than the 400 genes) take the n x n dual on the FP64 scaled rows whatever the
        # Gram precision, so they match the oracle's eigenvectors, not the split-precision
        # single run's.
        if n <= 256:
            np.testing.assert_allclose(out.variance[i], var_o, rtol=1e-9, err_msg=f"subset {i}")
            assert _angle(oracle, out.loadings[i], load_o, 9) < 1e-9, i
        else:
            assert _angle(oracle, out.loadings[i], o1.loadings, 9) < 1e-6, i
        # and the subset's own pipeline agrees on the oracle's variance
        if prec_gram == "fp64":
            np.testing.assert_allclose(out.variance[i], var_o, rtol=1e-9, err_msg=f"subset {i}")
            assert abs(out.total_variance[i] - tot_o) <= 1e-11 * tot_o
            # scores: (Z - mean Z) V with the oracle's loadings, through the oracle's
            # projection (FP64 accumulation, f32 output)
            Xo = np.empty((n, 30), np.float32)
            Xo[:] = -((colsum / n) @ load_o).astype(np.float32)
            oracle.project(*ctx, gmap, load_o, Xo, True, mean, sd, 10.0)
            sign = np.sign((X[a:b, :9] * Xo[:, :9]).sum(0))
            np.testing.assert_allclose(X[a:b, :9] * sign, Xo[:, :9], atol=1e-4 * np.abs(Xo[:, :9]).max())


def test_batch_tiles_ranges_and_eig_groups(batch, oracle):
    """Subsets cut across 1,024-row tiles, Gram row ranges of 256 rows (every subset
    above 256 rows goes through partial slots and the ordered reduce), eigensolves 3
    subsets per call: the same results as the default plan."""
    _, o_def, X_def, i_def, _ = _run(batch, prec_gram="fp64", prec_knn="fp32")
    _, o, X, idx, _ = _run(batch, prec_gram="fp64", prec_knn="fp32", tile_rows=1024, gram_rows=256, eig_batch=3)
    assert o.tile_rows == 1024 and o.gram_rows == 256
    np.testing.assert_allclose(o.variance, o_def.variance, rtol=1e-11)
    np.testing.assert_allclose(o.total_variance, o_def.total_variance, rtol=1e-12)
    for i in range(len(SIZES)):
        assert _angle(oracle, o.loadings[i], o_def.loadings[i], 9) < 1e-9
    np.testing.assert_allclose(X, X_def, rtol=1e-5, atol=1e-5)
    assert (np.sort(idx, 1) == np.sort(i_def, 1)).all(1).mean() > 0.999


def test_batch_sparse_fp64_gram_matches_dense(batch, oracle, monkeypatch):
    """The batch's FP64 Gram by sparse pair products (the default at FP64: 64-row
    groups per subset, one item per block and range, ranges of 4 groups with gram_rows
    = 256 so that every subset above 256 rows goes through partial slots) against the
    dense FP64 MFMA form (CREST_GRAM_SPARSE=0): both exact FP64 sums, so the
    eigenpairs agree to FP64 rounding."""
    _, o_sp, X_sp, i_sp, _ = _run(batch, prec_gram="fp64", prec_knn="fp32")
    _, o_sr, _, _, _ = _run(batch, prec_gram="fp64", prec_knn="fp32", gram_rows=256)
    monkeypatch.setenv("CREST_GRAM_SPARSE", "0")
    _, o_d, X_d, i_d, _ = _run(batch, prec_gram="fp64", prec_knn="fp32")
    for o in (o_sp, o_sr):
        np.testing.assert_allclose(o.variance, o_d.variance, rtol=1e-11)
        np.testing.assert_allclose(o.total_variance, o_d.total_variance, rtol=1e-12)
        for i in range(len(SIZES)):
            assert _angle(oracle, o.loadings[i], o_d.loadings[i], 9) < 1e-9, i
    np.testing.assert_allclose(X_sp, X_d, rtol=1e-5, atol=1e-5)
    assert (np.sort(i_sp, 1) == np.sort(i_d, 1)).all(1).mean() > 0.999


def test_batch_autotuner_reports_its_choice(batch):
    _, o1, _, _, _ = _run(batch, prec_gram="fp16", prec_knn="fp16", tile_rows=4096)
    assert o1.gram_rows in (2048, 4096, 8192, 16384, 1 << 20)
    _, o2, _, _, _ = _run(batch, prec_gram="fp16", prec_knn="fp16", tile_rows=4096)
    assert o2.gram_rows == o1.gram_rows and not o2.gram_tuned  # cached per shape


def test_batch_graph_replays(batch, oracle):
    """Captured graphs replayed several times, with eager batch runs and device
    synchronisations in between: every replay matches the eager run."""
    from crest_amd._lib import check, lib
    _, o_e, X_e, i_e, d_e = _run(batch, prec_gram="fp16", prec_knn="fp16")
    from crest_amd.pipeline import BatchPipeline, DeviceCSR
    indptr, indices, data, off = batch
    csr = DeviceCSR.upload(indptr, indices, data, G)
    pg = BatchPipeline(csr, off, np.arange(G, dtype=np.int32), n_comps=30, k=15, prec_gram="fp16",
                       prec_knn="fp16", use_graph=True)
    pe = BatchPipeline(csr, off, np.arange(G, dtype=np.int32), n_comps=30, k=15, prec_gram="fp16",
                       prec_knn="fp16")
    for rep in range(4):
        o = pg.run()
        assert o.graph_replayed == (rep > 0)
        X, (idx, dist) = pg.x_pca(), pg.knn()
        np.testing.assert_allclose(o.variance, o_e.variance, rtol=1e-9)
        np.testing.assert_allclose(X, X_e, rtol=1e-5, atol=1e-5)
        assert (np.sort(idx, 1) == np.sort(i_e, 1)).all(1).mean() > 0.999
        np.testing.assert_allclose(dist, d_e, rtol=1e-5, atol=1e-6)
        pe.run()  # eager work on the same stream between replays
        check(lib().crest_synchronize())
    # the eager object's own results are unchanged by the replays
    np.testing.assert_allclose(pe.x_pca(), X_e, rtol=1e-5, atol=1e-5)


# Subsets with fewer cells than genes: below 1,600 genes, the ones of at most 256, 512,
# 1,024 and 1,536 cells take the n x n dual eigensolve (one bucket each), the 2,000-cell
# one the d x d covariance.
G2 = 1600
SIZES2 = [51, 256, 300, 600, 1100, 1536, 2000]


@pytest.mark.parametrize("prec,sizes", [("fp64", SIZES2), ("fp16", SIZES2), ("fp64", [40, 300, 700, 1200])])
def test_batch_dual_eigensolve_matches_covariance(oracle, monkeypatch, prec, sizes):
    """The dual's variance, total and loadings against the covariance path of the same
    batch (CREST_BATCH_DUAL=0) and, in FP64, against the oracle's eigensolve of each
    subset's Gram; the last case has no subset on the covariance path at all."""
    from crest_amd.pipeline import BatchPipeline, DeviceCSR
    indptr, indices, data, off = synth.make_subsets(sizes, G2, 10, density=0.05, seed=5)
    csr = DeviceCSR.upload(indptr, indices, data, G2)
    res = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("CREST_BATCH_DUAL", mode)
        p = BatchPipeline(csr, off, np.arange(G2, dtype=np.int32), n_comps=30, k=15, prec_gram=prec,
                          prec_knn="fp32")
        o = p.run()
        res[mode] = (o, p.x_pca(), p.knn()[0])
    o, X, idx = res["1"]
    oc, Xc, ic = res["0"]
    # the Gram (and so the total) is the same; the dual works on the FP64 scaled rows,
    # the covariance path on the Gram of the chosen precision
    np.testing.assert_allclose(o.total_variance, oc.total_variance, rtol=1e-14)
    vr = 1e-10 if prec == "fp64" else 1e-6
    np.testing.assert_allclose(o.variance, oc.variance, rtol=vr)
    for i in range(len(sizes)):
        assert _angle(oracle, o.loadings[i], oc.loadings[i], 9) < (1e-9 if prec == "fp64" else 1e-6), i
        a, b = int(off[i]), int(off[i + 1])
        _check_knn(oracle, X[a:b], idx[a:b].astype(np.int64), b - a)
    if prec != "fp64":
        return
    gmap = np.arange(G2, dtype=np.int32)
    for i in sorted({0, 2, len(sizes) - 2}):  # subsets from different buckets
        a, b = int(off[i]), int(off[i + 1])
        n = b - a
        ip = indptr[a:b + 1] - indptr[a]
        ctx = (indices[indptr[a]:indptr[b]], data[indptr[a]:indptr[b]], np.arange(n, dtype=np.int64), gmap, 1e4,
               True, None, ip, 0)
        st = np.zeros((G2, 5))
        oracle.gene_stats(*ctx, st)
        mean = st[:, 0] / n
        sd = np.sqrt(np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0))
        sd[sd == 0] = 1
        gram, colsum = np.zeros((G2, G2)), np.zeros(G2)
        oracle.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=4)
        var_o, load_o, tot_o = oracle.pca_from_gram(gram, colsum, n, 30)
        np.testing.assert_allclose(o.variance[i], var_o, rtol=1e-9, err_msg=f"subset {i}")
        assert abs(o.total_variance[i] - tot_o) <= 1e-11 * tot_o
        assert _angle(oracle, o.loadings[i], load_o, 9) < 1e-9, i


@pytest.mark.parametrize("genes,sizes,seed", [(45, [40, 100, 300], 8), (G, SIZES, 21), (G2, SIZES2, 5)])
def test_batch_tridiagonalization_kernel_matches_rocsolver(oracle, monkeypatch, genes, sizes, seed):
    """The one-workgroup-per-matrix tridiagonalization (CREST_SYTRD=wg) against
    rocSOLVER's dsytrd (CREST_SYTRD=rocsolver) on the same batch: orders 45 (one tile),
    400 (a partial last panel), 1,600 and the dual orders 256 ... 1,536."""
    from crest_amd.pipeline import BatchPipeline, DeviceCSR
    indptr, indices, data, off = synth.make_subsets(sizes, genes, 10, density=0.07, seed=seed)
    csr = DeviceCSR.upload(indptr, indices, data, genes)
    res = {}
    monkeypatch.setenv("CREST_BATCH_EIG", "tri")  # the covariances through the tridiagonalization too
    for mode in ("rocsolver", "wg"):
        monkeypatch.setenv("CREST_SYTRD", mode)
        p = BatchPipeline(csr, off, np.arange(genes, dtype=np.int32), n_comps=30, k=15, prec_gram="fp64",
                          prec_knn="fp32")
        res[mode] = p.run()
    o, r = res["wg"], res["rocsolver"]
    assert np.isfinite(o.variance).all() and np.isfinite(o.loadings).all()
    np.testing.assert_allclose(o.total_variance, r.total_variance, rtol=1e-14)
    np.testing.assert_allclose(o.variance, r.variance, rtol=1e-10, atol=1e-12 * r.variance.max())
    for i in range(len(sizes)):
        assert _angle(oracle, o.loadings[i], r.loadings[i], 9) < 1e-9, i
        V = o.loadings[i]
        np.testing.assert_allclose(V.T @ V, np.eye(V.shape[1]), atol=1e-10)


# Orders from 512 up take the filtered subspace iteration, all covariances of a batch
# side by side; these subsets are all above the dual buckets below 640 genes.
G3 = 640
SIZES3 = [600, 900, 1500, 2200, 3100, 700, 1300, 4000]


@pytest.mark.parametrize("prec,pcs", [("fp64", 30), ("fp16", 30), ("fp64", 50)])
def test_batch_filtered_subspace_matches_tridiagonal(oracle, monkeypatch, capfd, prec, pcs):
    """The batched filtered subspace iteration (default from order 512) against the
    tridiagonal path of the same batch (CREST_BATCH_EIG=tri) and, in FP64, the oracle's
    eigensolve of a subset's Gram: every matrix solved in the batch, variances, totals,
    orthonormal loadings and the leading subspaces."""
    from crest_amd.pipeline import BatchPipeline, DeviceCSR
    indptr, indices, data, off = synth.make_subsets(SIZES3, G3, 10, density=0.06, seed=13)
    csr = DeviceCSR.upload(indptr, indices, data, G3)
    res = {}
    monkeypatch.setenv("CREST_CHEFSI_TRACE", "1")
    for mode in ("chefsi", "tri"):
        if mode == "tri":
            monkeypatch.setenv("CREST_BATCH_EIG", "tri")
        else:
            monkeypatch.delenv("CREST_BATCH_EIG", raising=False)
        print(f"[batch-chefsi] {prec} p={pcs} {mode}", flush=True)
        p = BatchPipeline(csr, off, np.arange(G3, dtype=np.int32), n_comps=pcs, k=15, prec_gram=prec,
                          prec_knn="fp32")
        res[mode] = (p.run(), p.x_pca())
        if mode == "chefsi":
            err = capfd.readouterr().err
            assert f"[chefsi-batch] {len(SIZES3)} matrices: {len(SIZES3)} solved" in err, err[-3000:]
    o, X = res["chefsi"]
    r, Xr = res["tri"]
    assert np.isfinite(o.variance).all() and np.isfinite(o.loadings).all()
    np.testing.assert_allclose(o.total_variance, r.total_variance, rtol=1e-14)
    np.testing.assert_allclose(o.variance, r.variance, rtol=1e-10, atol=1e-12 * r.variance.max())
    for i in range(len(SIZES3)):
        V = o.loadings[i]
        np.testing.assert_allclose(V.T @ V, np.eye(pcs), atol=1e-11)
        assert _angle(oracle, V, r.loadings[i], 9) < 1e-9, i
        a, b = int(off[i]), int(off[i + 1])
        sign = np.sign((X[a:b, :9] * Xr[a:b, :9]).sum(0))
        np.testing.assert_allclose(X[a:b, :9] * sign, Xr[a:b, :9], atol=1e-4 * np.abs(Xr[a:b, :9]).max())
    if prec != "fp64":
        return
    gmap = np.arange(G3, dtype=np.int32)
    for i in (0, 4, 7):
        a, b = int(off[i]), int(off[i + 1])
        n = b - a
        ip = indptr[a:b + 1] - indptr[a]
        ctx = (indices[indptr[a]:indptr[b]], data[indptr[a]:indptr[b]], np.arange(n, dtype=np.int64), gmap, 1e4,
               True, None, ip, 0)
        st = np.zeros((G3, 5))
        oracle.gene_stats(*ctx, st)
        mean = st[:, 0] / n
        sd = np.sqrt(np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0))
        sd[sd == 0] = 1
        gram, colsum = np.zeros((G3, G3)), np.zeros(G3)
        oracle.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=4)
        var_o, load_o, tot_o = oracle.pca_from_gram(gram, colsum, n, pcs)
        np.testing.assert_allclose(o.variance[i], var_o, rtol=1e-9, err_msg=f"subset {i}")
        assert abs(o.total_variance[i] - tot_o) <= 1e-11 * tot_o
        assert _angle(oracle, o.loadings[i], load_o, 9) < 1e-9, i


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_knn_certificate(batch, oracle, monkeypatch, use_graph):
    """The batch's kNN certificate, then with a bound scaled by 1e9 its recheck: every
    query of a full screen list rescreened in FP32 among its own subset's cells (the
    31-cell subset's lists are not full, so it stays certified), results still exact."""
    import crest_amd as ca
    n = int(batch[3][-1])
    _run(batch, prec_gram="fp64", prec_knn="fp16", use_graph=use_graph)
    c0 = ca.knn_certificate()
    assert c0["checked"] == n and c0["uncertified"] == 0, c0
    monkeypatch.setenv("CREST_KNN_CERT_SCALE", "1e9")
    pipe = _run(batch, prec_gram="fp64", prec_knn="fp16", use_graph=use_graph)[0]
    for rep in range(2):  # the second run of the same pipeline replays its graphs
        if rep:
            pipe.run()
        X, idx = pipe.x_pca(), pipe.knn()[0]
        c1 = ca.knn_certificate()
        assert c1["checked"] == n and c1["rechecked"] >= n - SIZES[0], c1
        assert c1["uncertified"] <= c1["rechecked"], c1
        off = batch[3]
        for i in range(len(SIZES)):
            a, b = int(off[i]), int(off[i + 1])
            _check_knn(oracle, X[a:b], idx[a:b].astype(np.int64), b - a)
