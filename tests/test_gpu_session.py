"""Device sessions (crest_session_*): gene statistics + seurat HVG as one device pass
against the oracle's statistics and its pandas restatement of the selection
(crest/pp.py:201-292), and the session form of tl.pca (crest/tl.py:37-91) over
filtered cells and genes and over each chunk source, against the oracle.

Bars: HVG index sets equal; the device's means / dispersions within 1e-13 relative
(its log / log1p against numpy's); PCA variances 1e-9 relative, loadings within
1e-8 rad; scores 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import synth  # noqa: E402
from crest_amd.session import Session  # noqa: E402


def oracle_stats(oracle, bf):
    st = np.zeros((bf.n_vars, 5))
    for ctx in bf.iter_ctx():
        oracle.gene_stats(*ctx, st)
    return st


def oracle_hvg(oracle, st, n, n_top, n_bins=20):
    m = st[:, 2] / n
    v = np.maximum((st[:, 3] - n * m * m) / (n - 1), 0.0)
    return oracle.hvg_seurat_pandas(m, v, n_top_genes=n_top, n_bins=n_bins)


def with_edge_genes(ip, ix, dv, n_genes):
    """Append genes that stress the selection: never expressed (zero mean -> 1e-12,
    NaN dispersion), expressed in exactly one cell, and exact copies of existing genes
    (equal dispersions: ties at the cut)."""
    n = len(ip) - 1
    rows = np.repeat(np.arange(n), np.diff(ip))
    extra_r, extra_g, extra_v = [], [], []
    g = n_genes
    g += 3  # three never-expressed genes
    extra_r.append(np.array([7]))
    extra_g.append(np.array([g]))
    extra_v.append(np.array([5.0]))
    g += 1
    for src in (0, 1, 2, 3):  # copies of genes 0..3
        sel = ix == src
        extra_r.append(rows[sel])
        extra_g.append(np.full(sel.sum(), g))
        extra_v.append(dv[sel])
        g += 1
    r = np.concatenate([rows] + extra_r)
    c = np.concatenate([ix.astype(np.int64)] + extra_g)
    v = np.concatenate([dv.astype(np.float32)] + [e.astype(np.float32) for e in extra_v])
    o = np.lexsort((c, r))
    r, c, v = r[o], c[o], v[o]
    ip2 = np.zeros(n + 1, np.int64)
    np.add.at(ip2, r + 1, 1)
    return np.cumsum(ip2), c.astype(np.uint32), v, g


@pytest.mark.parametrize("cells,genes,clusters,seed,n_top,n_bins", [
    (2000, 3000, 6, 5, 500, 20), (2000, 3000, 6, 5, None, 20), (1500, 1200, 4, 6, 300, 5),
    (3000, 5000, 10, 7, 1000, 50), (800, 700, 3, 8, 100, 1)])
def test_hvg_seurat_on_device(oracle, cells, genes, clusters, seed, n_top, n_bins):
    ip, ix, dv, _ = synth.make_counts(cells, genes, clusters, density=0.06, seed=seed)
    ip, ix, dv, G = with_edge_genes(ip, ix, dv, genes)
    bf = ca.BioFrame.from_csr(ip, ix, dv, G)
    bf.chunk_nnz = max(1, len(dv) // 4)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    print(f"[hvg] {cells} x {G}, n_top {n_top}, {n_bins} bins", flush=True)
    ca.pp.highly_variable_genes(bf, n_top_genes=n_top, n_bins=n_bins)
    st = oracle_stats(oracle, bf)
    hv, dn = oracle_hvg(oracle, st, bf.n_obs, n_top, n_bins)
    np.testing.assert_array_equal(np.asarray(bf.var["highly_variable"], bool), hv)
    np.testing.assert_allclose(bf.var["dispersions_norm"], dn, rtol=1e-9, atol=1e-12, equal_nan=True)
    m = st[:, 2] / bf.n_obs
    m[m == 0] = 1e-12
    np.testing.assert_allclose(bf.var["means"], np.log1p(m), rtol=1e-13)
    np.testing.assert_array_equal(bf.var["n_cells_expr"], st[:, 4].astype(np.int64))
    ml = st[:, 0] / bf.n_obs
    np.testing.assert_allclose(bf.var["mean_log"], ml, rtol=1e-12)
    if n_top is not None:
        assert int(hv.sum()) >= n_top  # ties at the cut are all kept


def test_hvg_all_genes_nan_window(oracle):
    """No n_top: the mean / dispersion window (NaN normalised dispersions count as 0)."""
    ip, ix, dv, _ = synth.make_counts(1000, 400, 4, density=0.1, seed=9)
    bf = ca.BioFrame.from_csr(ip, ix, dv, 400)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    ca.pp.highly_variable_genes(bf, min_mean=0.05, max_mean=2.0, min_disp=0.2)
    st = oracle_stats(oracle, bf)
    m = st[:, 2] / bf.n_obs
    v = np.maximum((st[:, 3] - bf.n_obs * m * m) / (bf.n_obs - 1), 0.0)
    hv, _ = oracle.hvg_seurat_pandas(m, v, min_mean=0.05, max_mean=2.0, min_disp=0.2)
    np.testing.assert_array_equal(np.asarray(bf.var["highly_variable"], bool), hv)


def _pca_oracle(oracle, bf, n_comps):
    mask = np.asarray(bf.var["highly_variable"], bool)
    d = int(mask.sum())
    sub = np.full(bf.n_vars, -1, np.int32)
    sub[mask] = np.arange(d, dtype=np.int32)
    st = oracle_stats(oracle, bf)
    n = bf.n_obs
    m = st[:, 0] / n
    sd = np.sqrt(np.maximum((st[:, 1] - n * m * m) / (n - 1), 0.0))
    sd[sd == 0] = 1.0
    mean, std = np.ascontiguousarray(m[mask]), np.ascontiguousarray(sd[mask])
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    for ctx in bf.iter_ctx():
        oracle.gram_accumulate(*ctx, sub, gram, colsum, True, mean, std, 10.0, threads=8)
    var, load, tot = oracle.pca_from_gram(gram, colsum, n, n_comps)
    X = np.empty((n, var.size), np.float32)
    X[:] = -((colsum / n) @ load).astype(np.float32)
    for ctx in bf.iter_ctx():
        oracle.project(*ctx, sub, load, X, True, mean, std, 10.0)
    return var, load, tot, X


@pytest.mark.parametrize("source", ["csr", "triplets", "parquet"])
def test_session_pca_over_chunk_sources(oracle, tmp_path, source):
    """tl.pca through the session with filtered cells and genes, scale statistics
    derived on the device (no HVG cache) and several chunks, from each store type."""
    ip, ix, dv, _ = synth.make_counts(3000, 800, 6, density=0.08, seed=12)
    bf = ca.BioFrame.from_csr(ip, ix, dv, 800)
    if source == "triplets":
        rows = np.repeat(np.arange(len(ip) - 1), np.diff(ip)).astype(np.uint32)
        bf = ca.BioFrame.from_triplets(rows, ix, dv, len(ip) - 1, 800)
    elif source == "parquet":
        bf.write_parquet(tmp_path / "pq", nnz_per_part=len(dv) // 5)
        bf = ca.read_parquet(tmp_path / "pq")
    bf.chunk_nnz = len(dv) // 6
    rng = np.random.default_rng(3)
    bf = bf.filter_cells(rng.random(bf.n_obs) > 0.07).filter_genes(rng.random(bf.n_vars) > 0.1)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    bf.var["highly_variable"] = rng.random(bf.n_vars) > 0.4
    ca.pp.scale(bf, max_value=10)
    print(f"[session] tl.pca over {source}", flush=True)
    ca.tl.pca(bf, n_comps=20)
    var, load, tot, X = _pca_oracle(oracle, bf, 20)
    u = bf.uns["pca"]
    np.testing.assert_allclose(u["variance"], var, rtol=1e-9)
    np.testing.assert_allclose(u["variance_ratio"], var / tot, rtol=1e-9)
    ev = var
    gaps = (ev[:-1] - ev[1:]) / ev[:-1]
    r = int(np.argmax(gaps[:8]) + 1)
    assert oracle.principal_angle(u["projection"]["loadings"][:, :r], load[:, :r]) < 1e-8
    np.testing.assert_allclose(np.abs(bf.obsm["X_pca"][:, :r]), np.abs(X[:, :r]), rtol=1e-5, atol=1e-5)
    assert bf.obsm["X_pca"].shape == (bf.n_obs, 20) and bf.varm["PCs"].shape == (bf.n_vars, 20)


def test_session_errors():
    s = Session(10, 5)
    try:
        with pytest.raises(ValueError, match="begin_gram first"):
            s.solve(3)
        ctx = (np.zeros(0, np.uint32), np.zeros(0, np.float32), np.arange(5, dtype=np.int64),
               np.arange(12, dtype=np.int32), 1e4, True, None, np.zeros(6, np.int64), 0)
        with pytest.raises(ValueError, match="n_vars"):
            s.push_stats(ctx)  # gene_map maps onto 12 genes, the session has 10
        with pytest.raises(ValueError, match="no statistics"):
            s.begin_gram(np.arange(10, dtype=np.int32), 10, True, None, None, 10.0)
    finally:
        s.close()
