"""The benchmark configurations at their full sizes through the GPU library, checked
against the oracle (tests/test_crest.py:54-118 restated at each configuration):

* 100,000 cells x 33,538 genes (20 clusters, seed 1) through the operator path
  (BioFrame chunks of 16,384 cells, seurat HVG 2,000, scale 10, tl.pca 50,
  knn_graph(X, 15)): the HVG index set equal to the oracle's under log1pf and under
  the correctly rounded log1p, PCA and sampled kNN against the oracle;
* 1,000,000 cells x 2,000 pre-selected HVGs (30 clusters, seed 2) through the fused
  pipeline in each precision: FP64 Gram, split-FP16, split-BF16 and FP32, against
  the oracle's FP64 Gram and eigensolve, kNN on 4,096 sampled queries;
* 256 subsets of log-uniform [200, 50,000] cells (10 clusters each, seed 3) in one
  batched call: every subset's total variance against the oracle, twelve subsets'
  PCA and kNN in full;
* 2,000,000 cells x 2,000 HVGs (seed 3) streamed from host memory in chunks of
  32,768 cells against the same CSR resident on the device and the oracle.

Every kNN search must also leave no query unproven by its certificate (the screen's
rounding bound, or the exact recheck of the queries it cannot prove).

The oracle is the C restatement of the reference (oracle/) with numpy LAPACK for the
eigensolve.  Each case prints its name before its GPU work."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import _native as N  # noqa: E402
from crest_amd import synth  # noqa: E402

THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))


def log(msg):
    print(f"[configs] {msg}", flush=True)


def ctx_rows(ip, ix, dv, cmap, gmap, lo, hi, target=1e4, log1p=True):
    """The reference's 9-tuple for cells lo..hi-1 of a CSR (crest/core.py:265-278)."""
    a, b = int(ip[lo]), int(ip[hi])
    return (ix[a:b], dv[a:b], cmap, gmap, target, log1p, None, np.ascontiguousarray(ip[lo:hi + 1]), lo)


def oracle_pca(oracle, ip, ix, dv, n_genes, sub, n_comps, lo=0, hi=None, step=131072, stats=None):
    """Oracle scale + PCA of cells lo..hi-1 (kernels.rs:141-160, pp.py:123-126,
    kernels.rs:221-348): gene statistics (unless given), mean/std, FP64 Gram, eigh."""
    hi = len(ip) - 1 if hi is None else hi
    n = hi - lo
    cmap = np.full(len(ip) - 1, -1, np.int64)
    cmap[lo:hi] = np.arange(n)
    gall = np.arange(n_genes, dtype=np.int32)
    if stats is None:
        stats = np.zeros((n_genes, 5))
        for a in range(lo, hi, step):
            oracle.gene_stats(*ctx_rows(ip, ix, dv, cmap, gall, a, min(hi, a + step)), stats)
    d = int(sub.max()) + 1
    hv = np.flatnonzero(sub >= 0)
    order = np.empty(d, np.int64)
    order[sub[hv]] = hv
    m = stats[:, 0] / n
    sd = np.sqrt(np.maximum((stats[:, 1] - n * m * m) / (n - 1), 0.0))
    sd[sd == 0] = 1.0
    mean, std = np.ascontiguousarray(m[order]), np.ascontiguousarray(sd[order])
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    for a in range(lo, hi, step):
        oracle.gram_accumulate(*ctx_rows(ip, ix, dv, cmap, gall, a, min(hi, a + step)), sub, gram, colsum, True,
                               mean, std, 10.0, threads=THREADS)
    var, load, tot = oracle.pca_from_gram(gram, colsum, n, n_comps)
    ev = np.linalg.eigvalsh(oracle.cov_from_gram(gram, colsum, n)[0])[::-1]
    return dict(stats=stats, mean=mean, std=std, var=var, load=load, tot=tot, ev=ev)


def signal_r(ev, p):
    """Size of the leading subspace at the largest relative eigengap among the first p."""
    rel = (ev[:-1] - ev[1:]) / ev[:-1]
    return int(np.argmax(rel[1:p - 1]) + 2)


def check_knn(oracle, X, idx, dist, queries, k):
    """Index sets equal the FP64 exact search except distance ties; distances equal the
    reference's f32 formula (knn.rs:48-66) bit for bit."""
    ref_i, ref_d2 = oracle.knn_exact_f64(X, k, queries, threads=THREADS)
    bad, ties = oracle.knn_mismatch_ties(idx[queries], X, queries, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(X, np.repeat(queries, k), idx[queries].ravel()).reshape(len(queries), k)
    np.testing.assert_array_equal(dist[queries], pd)
    return bad


# --------------------------------------------------------------------- 100k x 33,538
@pytest.fixture(scope="module")
def c2():
    ip, ix, dv, info = synth.make_counts(100_000, 33_538, 20, density=0.07, seed=1)
    log(f"100000 x 33538 generated, nnz {info['nnz']}")
    return ip, ix, dv


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_c2_operator_path(c2, oracle, prec):
    ip, ix, dv = c2
    n, G = len(ip) - 1, 33_538
    log(f"100000 x 33538 operator path, {prec} Gram and screen")
    prev = ca.get_precision()
    ca.set_precision(prec)
    try:
        bf = ca.BioFrame.from_csr(ip, ix, dv, G)
        bf.chunk_nnz = int(ip[16384])
        ca.pp.normalize_total(bf, 1e4)
        ca.pp.log1p(bf)
        ca.pp.highly_variable_genes(bf, n_top_genes=2000)
        ca.pp.scale(bf, max_value=10)
        ca.tl.pca(bf, n_comps=50)
        idx, dist = N.knn_graph(bf.obsm["X_pca"], 15)
        cert = ca.knn_certificate()
    finally:
        ca.set_precision(prev)
    # every query's set proven against the screen's bound or rechecked exactly
    assert cert["checked"] == n and cert["uncertified"] == 0, cert
    # HVG: the oracle's statistics under log1pf (the reference's rounding) and under
    # the correctly rounded log1p, each through the pandas seurat restatement
    cmap, gall = np.arange(n, dtype=np.int64), np.arange(G, dtype=np.int32)
    hv_gpu = np.asarray(bf.var["highly_variable"], bool)
    st_by_mode = {}
    for mode in (1, 0):
        oracle.set_log1p_mode(mode)
        try:
            st = np.zeros((G, 5))
            for a in range(0, n, 16384):
                oracle.gene_stats(*ctx_rows(ip, ix, dv, cmap, gall, a, min(n, a + 16384)), st)
        finally:
            oracle.set_log1p_mode(1)
        m = st[:, 2] / n
        v = np.maximum((st[:, 3] - n * m * m) / (n - 1), 0.0)
        hv, _ = oracle.hvg_seurat_pandas(m, v, n_top_genes=2000)
        np.testing.assert_array_equal(hv_gpu, hv, err_msg=f"log1p mode {mode}")
        st_by_mode[mode] = st
    assert hv_gpu.sum() == 2000
    sub = np.full(G, -1, np.int32)
    sub[hv_gpu] = np.arange(2000, dtype=np.int32)
    o = oracle_pca(oracle, ip, ix, dv, G, sub, 50, stats=st_by_mode[1])
    u = bf.uns["pca"]
    L = u["projection"]["loadings"]
    assert np.max(np.abs(u["variance_ratio"] - o["var"] / o["tot"])) <= 1e-7
    r = signal_r(o["ev"], 50)
    if prec == "fp64":
        np.testing.assert_allclose(u["variance"], o["var"], rtol=1e-9)
        assert oracle.principal_angle(L, o["load"]) <= 1e-6
    else:
        np.testing.assert_allclose(u["variance"], o["var"], rtol=1e-6)
        assert oracle.principal_angle(L[:, :r], o["load"][:, :r]) <= 1e-6
    q = np.random.default_rng(0).choice(n, 4096, replace=False)
    check_knn(oracle, bf.obsm["X_pca"], idx, dist, q, 15)


# --------------------------------------------------------------------- 1M x 2,000
@pytest.fixture(scope="module")
def c3(oracle):
    from crest_amd.pipeline import DeviceCSR
    ip, ix, dv, info = synth.make_counts(1_000_000, 2000, 30, density=0.08, seed=2)
    log(f"1000000 x 2000 generated, nnz {info['nnz']}")
    ref = oracle_pca(oracle, ip, ix, dv, 2000, np.arange(2000, dtype=np.int32), 50)
    log("1000000 x 2000 oracle PCA done")
    return DeviceCSR.upload(ip, ix, dv, 2000), ref


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec_gram,prec_knn", [("fp64", "fp16"), ("fp64", "bf16x3"), ("fp16", "fp16"),
                                                ("bf16x3", "bf16x3"), ("fp32", "fp32")])
def test_c3_pipeline(c3, oracle, prec_gram, prec_knn):
    """FP64 Gram: every one of the 50 components within 1e-6 rad and the variance
    ratio within 1e-7 (the accuracy bar).  The split and FP32 Grams: the 30-component
    signal subspace (relative eigengap 64%) within 1e-6 rad (split-BF16: 1e-5, its
    representation error ~2^-16) and the variance ratio within 1e-7."""
    from crest_amd.pipeline import Pipeline
    csr, ref = c3
    log(f"1000000 x 2000 pipeline, Gram {prec_gram}, screen {prec_knn}")
    p = Pipeline(csr, np.arange(2000, dtype=np.int32), n_comps=50, k=15, prec_gram=prec_gram, prec_knn=prec_knn)
    out = p.run()
    assert np.max(np.abs(out.variance / out.total_variance - ref["var"] / ref["tot"])) <= 1e-7
    r = signal_r(ref["ev"], 50)
    assert r == 30
    if prec_gram == "fp64":
        np.testing.assert_allclose(out.variance, ref["var"], rtol=1e-9)
        assert oracle.principal_angle(out.loadings, ref["load"]) <= 1e-6
    else:
        bar = 1e-5 if prec_gram == "bf16x3" else 1e-6
        assert oracle.principal_angle(out.loadings[:, :r], ref["load"][:, :r]) <= bar
    X = p.x_pca()
    idx, dist = p.knn()
    cert = ca.knn_certificate()
    assert cert["checked"] == len(X) and cert["uncertified"] == 0, cert
    q = np.random.default_rng(1).choice(len(X), 4096, replace=False)
    check_knn(oracle, X, idx, dist, q, 15)


# --------------------------------------------------------------------- 256 subsets
@pytest.fixture(scope="module")
def c4():
    from crest_amd.pipeline import DeviceCSR
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(200), np.log(50000), 256)).astype(np.int64)
    ip, ix, dv, off = synth.make_subsets(sizes, 2000, 10, density=0.07, seed=3)
    log(f"256 subsets generated, {int(off[-1])} cells, nnz {int(ip[-1])}")
    return ip, ix, dv, off, DeviceCSR.upload(ip, ix, dv, 2000)


@pytest.mark.timeout(900)
def test_c4_batch(c4, oracle):
    """One batched call (FP64 Gram, FP16 screen, graphs captured and replayed) over all
    256 subsets: every subset's total variance against the oracle's covariance trace
    (1e-10); twelve subsets across the size range in full: variances (1e-9), all 30
    loadings within 1e-6 rad, kNN of 512 sampled cells among the subset's own cells."""
    from crest_amd.pipeline import BatchPipeline
    ip, ix, dv, off, csr = c4
    hvg = np.arange(2000, dtype=np.int32)
    log("256 subsets batched, fp64 Gram, fp16 screen, graph capture")
    bp = BatchPipeline(csr, off, hvg, n_comps=30, k=15, prec_gram="fp64", prec_knn="fp16", use_graph=True)
    bp.run()
    out = bp.run()
    assert out.graph_replayed
    X = bp.x_pca()
    idx, dist = bp.knn()
    n_all = int(off[-1])
    cert = ca.knn_certificate()
    assert cert["checked"] == n_all and cert["uncertified"] == 0, cert
    gall = np.arange(2000, dtype=np.int32)
    for i in range(256):
        lo, hi = int(off[i]), int(off[i + 1])
        n = hi - lo
        cmap = np.full(n_all, -1, np.int64)
        cmap[lo:hi] = np.arange(n)
        st = np.zeros((2000, 5))
        oracle.gene_stats(*ctx_rows(ip, ix, dv, cmap, gall, lo, hi), st)
        m = st[:, 0] / n
        sd = np.sqrt(np.maximum((st[:, 1] - n * m * m) / (n - 1), 0.0))
        sd[sd == 0] = 1.0
        s1, s2 = np.zeros(2000), np.zeros(2000)
        oracle.scaled_moments(*ctx_rows(ip, ix, dv, cmap, gall, lo, hi), gall, s1, s2, True, m, sd, 10.0)
        tot = float(np.sum((s2 - s1 * s1 / n) / (n - 1)))
        assert abs(out.total_variance[i] - tot) <= 1e-10 * tot, (i, out.total_variance[i], tot)
    sizes = np.diff(off)
    pick = np.argsort(sizes)[np.linspace(0, 255, 12).astype(int)]
    for i in pick:
        lo, hi = int(off[i]), int(off[i + 1])
        log(f"subset {i}: {hi - lo} cells against the oracle")
        o = oracle_pca(oracle, ip, ix, dv, 2000, hvg, 30, lo=lo, hi=hi)
        np.testing.assert_allclose(out.variance[i], o["var"], rtol=1e-9, atol=1e-12 * o["tot"])
        assert oracle.principal_angle(out.loadings[i], o["load"]) <= 1e-6
        Xi = np.ascontiguousarray(X[lo:hi])
        q = np.random.default_rng(int(i)).choice(hi - lo, min(512, hi - lo), replace=False)
        check_knn(oracle, Xi, idx[lo:hi], dist[lo:hi], q, 15)


# --------------------------------------------------------------------- 2M x 2,000 streamed
@pytest.mark.timeout(900)
def test_c5_streamed(oracle):
    """2,000,000 cells streamed from pinned host memory in chunks of 32,768 cells
    (device memory: two chunk slots, the Gram and the kNN operands) against the same
    CSR resident on the device and against the oracle: variances equal to 1e-12, the
    50 loadings within 1e-6 rad of the oracle's, scores to 1e-5, and the kNN sets of
    4,096 sampled cells exact."""
    from crest_amd.pipeline import DeviceCSR, HostCSR, Pipeline
    ip, ix, dv, info = synth.make_counts(2_000_000, 2000, 30, density=0.08, seed=3)
    log(f"2000000 x 2000 generated, nnz {info['nnz']}")
    hvg = np.arange(2000, dtype=np.int32)
    host = HostCSR.from_arrays(ip, ix, dv, 2000, pinned=True)
    log("2000000 x 2000 streamed, chunks of 32768, fp64 Gram, fp16 screen")
    ps = Pipeline(host, hvg, n_comps=50, k=15, prec_gram="fp64", prec_knn="fp16", chunk_rows=32768,
                  stream_rows=32768)
    os_ = ps.run()
    Xs = ps.x_pca()
    i_s, d_s = ps.knn()
    cert = ca.knn_certificate()
    assert cert["checked"] == 2_000_000 and cert["uncertified"] == 0, cert
    del ps
    log("2000000 x 2000 resident, fp64 Gram, fp16 screen")
    pr = Pipeline(DeviceCSR.upload(ip, ix, dv, 2000), hvg, n_comps=50, k=15, prec_gram="fp64", prec_knn="fp16")
    orr = pr.run()
    Xr = pr.x_pca()
    np.testing.assert_allclose(os_.variance, orr.variance, rtol=1e-12)
    assert oracle.principal_angle(os_.loadings, orr.loadings) <= 1e-9
    np.testing.assert_allclose(np.abs(Xs), np.abs(Xr), rtol=1e-5, atol=1e-5)
    del pr
    ref = oracle_pca(oracle, ip, ix, dv, 2000, hvg, 50)
    np.testing.assert_allclose(os_.variance, ref["var"], rtol=1e-9)
    assert oracle.principal_angle(os_.loadings, ref["load"]) <= 1e-6
    q = np.random.default_rng(2).choice(len(Xs), 4096, replace=False)
    check_knn(oracle, Xs, i_s, d_s, q, 15)
