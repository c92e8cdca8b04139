"""Writes tests/golden/c1.npz: the oracle's outputs for the small plain-method
configuration -- 3,000 cells x 13,714 genes, 8 clusters, ~6% density, seed 0;
normalize_total 1e4, log1p, seurat HVG n_top 2,000, scale max_value 10, PCA 50,
knn_graph(X, 15), neighbors n_neighbors 15, QC on the raw counts -- on seeded
synthetic counts.  The gene statistics and the HVG index set are also stored under
the correctly rounded log1p (oracle mode 0) next to the reference's log1pf (mode 1,
the default).

Everything here is the oracle (oracle/, the C and numpy restatement of the
reference), never the product; the product's tests compare against the file.
Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from crest_amd import synth  # noqa: E402  (the seeded input generator)
from oracle import oracle as O  # noqa: E402

PARAMS = dict(n_cells=3000, n_genes=13714, n_clusters=8, density=0.06, seed=0)
N_TOP, N_COMPS, N_NEIGHBORS, K_GRAPH, MAX_VALUE = 2000, 50, 15, 15, 10.0
CHUNK_ROWS = 700


def inputs():
    p = PARAMS
    indptr, indices, data, _ = synth.make_counts(p["n_cells"], p["n_genes"], p["n_clusters"], density=p["density"],
                                                 seed=p["seed"])
    return indptr, indices, data


def csr_digest(indptr, indices, data) -> str:
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(indptr, np.int64), np.ascontiguousarray(indices, np.uint32),
              np.ascontiguousarray(data, np.float32)):
        h.update(a.tobytes())
    return h.hexdigest()


def ctx(indptr, indices, data, cmap, gmap, target_sum, log1p, lo, hi):
    a, b = int(indptr[lo]), int(indptr[hi])
    return (np.ascontiguousarray(indices[a:b], np.uint32), np.ascontiguousarray(data[a:b], np.float32), cmap, gmap,
            target_sum, log1p, None, np.ascontiguousarray(indptr[lo:hi + 1], np.int64), lo)


def oracle_outputs(indptr, indices, data):
    n, g = len(indptr) - 1, PARAMS["n_genes"]
    cmap = np.arange(n, dtype=np.int64)
    gmap = np.arange(g, dtype=np.int32)
    chunks = [(lo, min(n, lo + CHUNK_ROWS)) for lo in range(0, n, CHUNK_ROWS)]
    # QC on raw counts (pp.py:25-47), flag: every 50th gene
    flag = (np.arange(g) % 50 == 0).astype(np.uint8)
    qc = [np.zeros(n), np.zeros(n, np.uint32), np.zeros(n), np.zeros(g, np.uint32), np.zeros(g)]
    for lo, hi in chunks:
        O.qc(*ctx(indptr, indices, data, cmap, gmap, 0.0, False, lo, hi), flag, *qc)
    # gene statistics of the log-normalised matrix, seurat HVG (pandas restatement),
    # under log1pf (mode 1) and the correctly rounded log1p (mode 0)
    st_modes, hv_modes = {}, {}
    for mode in (0, 1):
        O.set_log1p_mode(mode)
        try:
            st_m = np.zeros((g, 5))
            for lo, hi in chunks:
                O.gene_stats(*ctx(indptr, indices, data, cmap, gmap, 1e4, True, lo, hi), st_m)
        finally:
            O.set_log1p_mode(1)
        m = st_m[:, 2] / n
        v = np.maximum((st_m[:, 3] - n * m * m) / (n - 1), 0.0)
        st_modes[mode], hv_modes[mode] = st_m, O.hvg_seurat_pandas(m, v, n_top_genes=N_TOP)[0]
    st, hv = st_modes[1], hv_modes[1]
    ml = st[:, 0] / n
    vl = np.maximum((st[:, 1] - n * ml * ml) / (n - 1), 0.0)
    sd = np.sqrt(vl)
    sd[sd == 0] = 1.0
    d = int(hv.sum())
    sub = np.full(g, -1, np.int32)
    sub[hv] = np.arange(d, dtype=np.int32)
    mean_h, sd_h = np.ascontiguousarray(ml[hv]), np.ascontiguousarray(sd[hv])
    # scaled Gram PCA and projection (tl.py:37-91)
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    for lo, hi in chunks:
        O.gram_accumulate(*ctx(indptr, indices, data, cmap, gmap, 1e4, True, lo, hi), sub, gram, colsum, True,
                          mean_h, sd_h, MAX_VALUE)
    var, load, tot = O.pca_from_gram(gram, colsum, n, N_COMPS)
    shift = (colsum / n) @ load
    X = np.empty((n, var.size), np.float32)
    X[:] = -shift.astype(np.float32)
    for lo, hi in chunks:
        O.project(*ctx(indptr, indices, data, cmap, gmap, 1e4, True, lo, hi), sub, load, X, True, mean_h, sd_h,
                  MAX_VALUE)
    # knn_graph(X, 15) (the configuration's k) and neighbours (pp.py:296-315):
    # knn_graph(X, n_neighbors - 1) + connectivities
    gidx, gdist = O.knn(X, K_GRAPH, threads=8)
    idx, dist = O.knn(X, N_NEIGHBORS - 1, threads=8)
    rows, cols, w = O.connectivities(idx, dist, N_NEIGHBORS)
    return dict(qc_flag=flag, qc_cell_total=qc[0], qc_cell_ngenes=qc[1], qc_cell_flag=qc[2], qc_gene_ncells=qc[3],
                qc_gene_total=qc[4], gene_stats=st, highly_variable=hv,
                highly_variable_cr=hv_modes[0], variance=var, total_variance=np.float64(tot), loadings=load, x_pca=X,
                graph_idx=gidx, graph_dist=gdist, knn_idx=idx, knn_dist=dist, conn_rows=rows, conn_cols=cols,
                conn_w=w)


def main():
    indptr, indices, data = inputs()
    out = oracle_outputs(indptr, indices, data)
    out["csr_sha256"] = np.array(csr_digest(indptr, indices, data))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
