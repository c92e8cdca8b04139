"""The fused device pipeline (crest_pipeline_run) against the oracle and against the
per-chunk operator path, in each precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import synth  # noqa: E402


@pytest.fixture(scope="module")
def data():
    indptr, indices, data, _ = synth.make_counts(20000, 600, 12, density=0.08, seed=5)
    return indptr, indices, data, 600


@pytest.fixture(scope="module")
def reference(data, oracle):
    """Oracle: streamed FP64 Gram PCA of all 600 genes, scaled, max_value 10."""
    indptr, indices, vals, g = data
    bf = ca.BioFrame.from_csr(indptr, indices, vals, g)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    st = np.zeros((g, 5))
    for ctx in bf.iter_ctx():
        oracle.gene_stats(*ctx, st)
    n = bf.n_obs
    mean = st[:, 0] / n
    var = np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0)
    sd = np.sqrt(var)
    sd[sd == 0] = 1
    gmap = np.arange(g, dtype=np.int32)
    gram, colsum = np.zeros((g, g)), np.zeros(g)
    for ctx in bf.iter_ctx():
        oracle.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=8)
    var_o, load_o, tot_o = oracle.pca_from_gram(gram, colsum, n, 30)
    shift = (colsum / n) @ load_o
    X = np.empty((n, 30), np.float32)
    X[:] = -shift.astype(np.float32)
    for ctx in bf.iter_ctx():
        oracle.project(*ctx, gmap, load_o, X, True, mean, sd, 10.0)
    return var_o, load_o, tot_o, X


@pytest.mark.parametrize("prec_gram,prec_knn,var_rtol,tot_rtol", [("fp64", "fp32", 1e-9, 1e-9),
                                                                  ("fp32", "fp32", 1e-5, 1e-7),
                                                                  ("bf16x3", "bf16x3", 1e-4, 1e-6),
                                                                  ("bf16x3", "fp16", 1e-4, 1e-6),
                                                                  ("fp16", "fp16", 1e-5, 1e-7)])
def test_pipeline(data, reference, oracle, prec_gram, prec_knn, var_rtol, tot_rtol):
    from crest_amd.pipeline import DeviceCSR, Pipeline
    indptr, indices, vals, g = data
    csr = DeviceCSR.upload(indptr, indices, vals, g)
    p = Pipeline(csr, np.arange(g, dtype=np.int32), n_comps=30, k=15, prec_gram=prec_gram, prec_knn=prec_knn,
                 chunk_rows=4096)
    out = p.run()
    var_o, load_o, tot_o, X_o = reference
    np.testing.assert_allclose(out.variance, var_o, rtol=var_rtol)
    assert abs(out.total_variance - tot_o) <= tot_rtol * tot_o
    lead = 10  # well separated leading components
    assert oracle.principal_angle(out.loadings[:, :lead], load_o[:, :lead]) < (1e-7 if prec_gram == "fp64" else 1e-3)
    X = p.x_pca()
    if prec_gram == "fp64":
        # same loadings up to eigen-solver rounding: scores agree closely
        np.testing.assert_allclose(np.abs(X[:, :lead]), np.abs(X_o[:, :lead]), rtol=1e-4, atol=1e-3)
    idx, dist = p.knn()
    q = np.random.default_rng(0).choice(len(X), 400, replace=False)
    ref_i, ref_d2 = oracle.knn_exact_f64(X, 15, q)
    bad, ties = oracle.knn_mismatch_ties(idx[q], X, q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(X, np.repeat(q, 15), idx[q].ravel()).reshape(len(q), 15)
    np.testing.assert_array_equal(dist[q], pd)
    for k_, v in out.times_ms.items():
        assert v >= 0.0, k_


@pytest.mark.parametrize("pinned,stream_rows", [(True, 4096), (False, 3000), (True, 50000)])
def test_streamed_matches_resident(data, reference, oracle, pinned, stream_rows):
    """The out-of-core form (host CSR streamed through two device slots) computes the
    same PCA and kNN as the device-resident pipeline: the Gram tiles follow the
    stream's chunk boundaries, so FP64 sums differ only in order."""
    from crest_amd.pipeline import DeviceCSR, HostCSR, Pipeline
    indptr, indices, vals, g = data
    hvg = np.arange(g, dtype=np.int32)
    res = Pipeline(DeviceCSR.upload(indptr, indices, vals, g), hvg, n_comps=30, k=15, prec_gram="fp64",
                   prec_knn="fp16", chunk_rows=4096)
    out_r = res.run()
    host = HostCSR.from_arrays(indptr, indices, vals, g, pinned=pinned)
    st = Pipeline(host, hvg, n_comps=30, k=15, prec_gram="fp64", prec_knn="fp16", chunk_rows=4096,
                  stream_rows=stream_rows)
    out_s = st.run()
    np.testing.assert_allclose(out_s.variance, out_r.variance, rtol=1e-11)
    assert abs(out_s.total_variance - out_r.total_variance) <= 1e-11 * out_r.total_variance
    assert oracle.principal_angle(out_s.loadings[:, :10], out_r.loadings[:, :10]) < 1e-8
    var_o, load_o, tot_o, X_o = reference
    np.testing.assert_allclose(out_s.variance, var_o, rtol=1e-9)
    X = st.x_pca()
    np.testing.assert_allclose(np.abs(X[:, :10]), np.abs(res.x_pca()[:, :10]), rtol=1e-4, atol=1e-4)
    idx, dist = st.knn()
    q = np.random.default_rng(1).choice(len(X), 300, replace=False)
    ref_i, ref_d2 = oracle.knn_exact_f64(X, 15, q)
    bad, ties = oracle.knn_mismatch_ties(idx[q], X, q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    # a second run reuses the slots and events
    out_s2 = st.run()
    np.testing.assert_allclose(out_s2.variance, out_s.variance, rtol=1e-12)


def test_streamed_errors():
    from crest_amd._lib import PipelineArgs, PipelineResult, lib
    import ctypes as C
    a = PipelineArgs()
    r = PipelineResult()
    v = np.zeros(4)
    assert lib().crest_pipeline_run_streamed(C.byref(a), None, None, None, 0, 0, v.ctypes.data, v.ctypes.data,
                                              C.byref(r)) != 0


def _oracle_pca(oracle, indptr, indices, vals, g, n_comps):
    n = len(indptr) - 1
    cmap, gmap = np.arange(n, dtype=np.int64), np.arange(g, dtype=np.int32)
    ctx = (indices, vals, cmap, gmap, 1e4, True, None, np.asarray(indptr, np.int64), 0)
    st = np.zeros((g, 5))
    oracle.gene_stats(*ctx, st)
    mean = st[:, 0] / n
    sd = np.sqrt(np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0))
    sd[sd == 0] = 1
    gram, colsum = np.zeros((g, g)), np.zeros(g)
    oracle.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=4)
    return oracle.pca_from_gram(gram, colsum, n, n_comps)


def test_many_subsets(oracle):
    """The batching configuration as a parity case: independent subsets of log-uniform
    sizes, each with its own generator and genes, through the same pipeline object
    types one after another (sizes below one workgroup's 512 queries included)."""
    from crest_amd.pipeline import DeviceCSR, Pipeline
    rng = np.random.default_rng(11)
    sizes = np.exp(rng.uniform(np.log(200), np.log(6000), 6)).astype(int)
    for i, n in enumerate(sizes):
        g = int(rng.integers(250, 700))
        indptr, indices, vals, _ = synth.make_counts(int(n), g, 10, density=0.07, seed=100 + i)
        p = Pipeline(DeviceCSR.upload(indptr, indices, vals, g), np.arange(g, dtype=np.int32), n_comps=30, k=15,
                     prec_gram="fp16", prec_knn="fp16", chunk_rows=2048)
        out = p.run()
        var_o, load_o, tot_o = _oracle_pca(oracle, indptr, indices, vals, g, 30)
        np.testing.assert_allclose(out.variance, var_o, rtol=1e-5, err_msg=f"subset {i} n={n} g={g}")
        X = p.x_pca()
        idx, dist = p.knn()
        q = np.arange(min(int(n), 300))
        ref_i, ref_d2 = oracle.knn_exact_f64(X, 15, q)
        bad, ties = oracle.knn_mismatch_ties(idx[q], X, q, ref_i, ref_d2)
        assert bad == ties, (i, n, bad, ties)


def test_pipeline_wide_scores(oracle):
    """150 components (wider than the FP32 register screen's 134 columns): the kNN of
    the fused pipeline runs the GEMM screen and exact rerank; sets equal the FP64
    exact search except ties, distances the reference's f32 formula."""
    from crest_amd.pipeline import DeviceCSR, Pipeline
    ip, ix, dv, _ = synth.make_counts(6000, 400, 8, density=0.1, seed=21)
    p = Pipeline(DeviceCSR.upload(ip, ix, dv, 400), np.arange(400, dtype=np.int32), n_comps=150, k=15,
                 prec_gram="fp64", prec_knn="fp64")
    out = p.run()
    assert out.variance.shape == (150,)
    X = p.x_pca()
    idx, dist = p.knn()
    q = np.arange(0, len(X), 7)
    ref_i, ref_d2 = oracle.knn_exact_f64(X, 15, q)
    bad, ties = oracle.knn_mismatch_ties(idx[q], X, q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(X, np.repeat(q, 15), idx[q].ravel()).reshape(len(q), 15)
    np.testing.assert_array_equal(dist[q], pd)
