"""The product's user-facing path on the golden C1-shaped run: calculate_qc_metrics,
normalize_total, log1p, highly_variable_genes (seurat, 300), scale(10), tl.pca(30),
neighbors(15), all through the GPU library, against the oracle's fixture."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as MG  # noqa: E402


@pytest.fixture(scope="module")
def run():
    z = np.load(os.path.join(HERE, "golden", "c1_small.npz"), allow_pickle=False)
    indptr, indices, data = MG.inputs()
    bf = ca.BioFrame.from_csr(indptr, indices, data, MG.PARAMS["n_genes"])
    bf.chunk_nnz = 20000  # several chunks
    bf.var["gene_name"] = np.array([("MT-%d" % i) if i % 50 == 0 else ("G%d" % i) for i in range(bf.n_vars)])
    obs, var = ca.pp.calculate_qc_metrics(bf, inplace=False)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    ca.pp.highly_variable_genes(bf, n_top_genes=MG.N_TOP)
    ca.pp.scale(bf, max_value=MG.MAX_VALUE)
    ca.tl.pca(bf, n_comps=MG.N_COMPS)
    ca.pp.neighbors(bf, n_neighbors=MG.N_NEIGHBORS)
    return z, bf, obs, var


def test_qc_bit_exact(run):
    z, bf, obs, var = run
    np.testing.assert_array_equal(obs["total_counts"], z["qc_cell_total"])
    np.testing.assert_array_equal(obs["n_genes_by_counts"], z["qc_cell_ngenes"].astype(np.int64))
    pct = np.divide(z["qc_cell_flag"], z["qc_cell_total"], out=np.zeros(len(obs["total_counts"])),
                    where=z["qc_cell_total"] > 0) * 100.0
    np.testing.assert_array_equal(obs["pct_counts_mt"], pct)
    np.testing.assert_array_equal(var["n_cells_by_counts"], z["qc_gene_ncells"].astype(np.int64))
    np.testing.assert_array_equal(var["total_counts"], z["qc_gene_total"])


def test_hvg_index_set_bit_exact(run):
    z, bf, _, _ = run
    np.testing.assert_array_equal(bf.var["highly_variable"].astype(bool), z["highly_variable"])


def test_pca_matches_fixture(run, oracle):
    z, bf, _, _ = run
    u = bf.uns["pca"]
    np.testing.assert_allclose(u["variance"], z["variance"], rtol=1e-9)
    np.testing.assert_allclose(u["variance_ratio"], z["variance"] / float(z["total_variance"]), rtol=1e-9)
    L = u["projection"]["loadings"]
    ev = z["variance"]
    # leading components separated by a clear gap: subspace angle to 1e-7 rad
    gaps = (ev[:-1] - ev[1:]) / ev[:-1]
    r = int(np.argmax(gaps[:10]) + 1)
    assert oracle.principal_angle(L[:, :r], z["loadings"][:, :r]) < 1e-7
    X = bf.obsm["X_pca"]
    np.testing.assert_allclose(np.abs(X[:, :r]), np.abs(z["x_pca"][:, :r]), rtol=1e-5, atol=1e-5)


def test_neighbors_match_fixture(run, oracle):
    z, bf, _, _ = run
    nb = bf.uns["neighbors"]
    idx, dist = nb["indices"], nb["distances"]
    X = bf.obsm["X_pca"]
    q = np.arange(len(X))
    ref_i, ref_d2 = oracle.knn_exact_f64(X, idx.shape[1], q)
    bad, ties = oracle.knn_mismatch_ties(idx, X, q, ref_i, ref_d2)
    assert bad == ties, (bad, ties)
    # against the fixture's own neighbour sets (computed on the oracle's scores)
    agree = oracle.knn_set_agreement(idx, z["knn_idx"])
    assert agree >= 0.99, agree
    # connectivities: the same edge set where the neighbour sets agree
    e_got = set(zip(nb["rows"].tolist(), nb["cols"].tolist()))
    e_ref = set(zip(z["conn_rows"].tolist(), z["conn_cols"].tolist()))
    assert len(e_got & e_ref) >= 0.98 * len(e_ref)
