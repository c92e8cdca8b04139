"""The product's user-facing path on the small plain-method configuration (3,000
cells x 13,714 genes, 8 clusters, seed 0): calculate_qc_metrics, normalize_total,
log1p, highly_variable_genes (seurat, 2,000), scale(10), tl.pca(50),
knn_graph(X, 15) and neighbors(15), all through the GPU library, against the
oracle's fixture tests/golden/c1.npz (tests/test_crest.py:54-118 restated at this
configuration)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import _native as N  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as MG  # noqa: E402


@pytest.fixture(scope="module")
def run():
    z = np.load(os.path.join(HERE, "golden", "c1.npz"), allow_pickle=False)
    indptr, indices, data = MG.inputs()
    assert MG.csr_digest(indptr, indices, data) == str(z["csr_sha256"])
    print(f"[c1] {len(indptr) - 1} x {MG.PARAMS['n_genes']}, nnz {len(data)}", flush=True)
    bf = ca.BioFrame.from_csr(indptr, indices, data, MG.PARAMS["n_genes"])
    bf.chunk_nnz = int(indptr[MG.CHUNK_ROWS])  # several chunks
    bf.var["gene_name"] = np.array([("MT-%d" % i) if i % 50 == 0 else ("G%d" % i) for i in range(bf.n_vars)])
    obs, var = ca.pp.calculate_qc_metrics(bf, inplace=False)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    stats = ca.pp.gene_stats(bf)
    ca.pp.highly_variable_genes(bf, n_top_genes=MG.N_TOP)
    ca.pp.scale(bf, max_value=MG.MAX_VALUE)
    ca.tl.pca(bf, n_comps=MG.N_COMPS)
    graph = N.knn_graph(bf.obsm["X_pca"], MG.K_GRAPH)
    ca.pp.neighbors(bf, n_neighbors=MG.N_NEIGHBORS)
    return z, bf, obs, var, stats, graph


def test_qc_bit_exact(run):
    z, bf, obs, var, _, _ = run
    np.testing.assert_array_equal(obs["total_counts"], z["qc_cell_total"])
    np.testing.assert_array_equal(obs["n_genes_by_counts"], z["qc_cell_ngenes"].astype(np.int64))
    pct = np.divide(z["qc_cell_flag"], z["qc_cell_total"], out=np.zeros(len(obs["total_counts"])),
                    where=z["qc_cell_total"] > 0) * 100.0
    np.testing.assert_array_equal(obs["pct_counts_mt"], pct)
    np.testing.assert_array_equal(var["n_cells_by_counts"], z["qc_gene_ncells"].astype(np.int64))
    np.testing.assert_array_equal(var["total_counts"], z["qc_gene_total"])


def test_gene_stats(run):
    """Counts exact; FP64 sums of log1pf values (the reference's rounding) to 1e-12."""
    z, _, _, _, st, _ = run
    np.testing.assert_array_equal(st[:, 4], z["gene_stats"][:, 4])
    np.testing.assert_allclose(st[:, :4], z["gene_stats"][:, :4], rtol=1e-12, atol=1e-12)


def test_hvg_index_set_bit_exact(run):
    """The seurat index set equals the oracle's under log1pf and under the correctly
    rounded log1p (the set does not depend on the last-ulp rounding here)."""
    z, bf, _, _, _, _ = run
    hv = np.asarray(bf.var["highly_variable"], bool)
    assert hv.sum() == MG.N_TOP
    np.testing.assert_array_equal(hv, z["highly_variable"])
    np.testing.assert_array_equal(hv, z["highly_variable_cr"])


def test_pca_matches_fixture(run, oracle):
    """Variance and variance ratio to 1e-9 relative / 1e-7 absolute; the 50-component
    loading subspace within 1e-6 rad; scores (up to sign) to 1e-4."""
    z, bf, _, _, _, _ = run
    u = bf.uns["pca"]
    np.testing.assert_allclose(u["variance"], z["variance"], rtol=1e-9)
    ratio = z["variance"] / float(z["total_variance"])
    assert np.max(np.abs(u["variance_ratio"] - ratio)) <= 1e-7
    L = u["projection"]["loadings"]
    assert oracle.principal_angle(L, z["loadings"]) <= 1e-6
    X = bf.obsm["X_pca"]
    np.testing.assert_allclose(np.abs(X), np.abs(z["x_pca"]), rtol=1e-5, atol=1e-4)


def test_knn_graph_k15(run, oracle):
    """knn_graph(X, 15): index sets equal the FP64 exact search on the GPU's scores
    except distance ties, distances equal the reference's f32 formula bit for bit, and
    the sets agree with the fixture's (computed on the oracle's scores)."""
    z, bf, _, _, _, (idx, dist) = run
    X = bf.obsm["X_pca"]
    q = np.arange(len(X))
    ref_i, ref_d2 = oracle.knn_exact_f64(X, MG.K_GRAPH, q)
    bad, ties = oracle.knn_mismatch_ties(idx, X, q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    pd = oracle.pair_dist(X, np.repeat(q, MG.K_GRAPH), idx.ravel()).reshape(idx.shape)
    np.testing.assert_array_equal(dist, pd)
    assert oracle.knn_set_agreement(idx, z["graph_idx"]) >= 0.999


def test_neighbors_match_fixture(run, oracle):
    z, bf, _, _, _, _ = run
    nb = bf.uns["neighbors"]
    idx = nb["indices"]
    assert idx.shape[1] == MG.N_NEIGHBORS - 1
    assert oracle.knn_set_agreement(idx, z["knn_idx"]) >= 0.999
    e_got = set(zip(nb["rows"].tolist(), nb["cols"].tolist()))
    e_ref = set(zip(z["conn_rows"].tolist(), z["conn_cols"].tolist()))
    assert len(e_got & e_ref) >= 0.995 * len(e_ref)
