"""Drop-in for the hot-path functions of the reference's native module `crest.crest`
(src/py.rs): same names, positional signatures, accumulation semantics and
exception classes, computed by libcrest_hip.so on the GPU.

Each chunk function takes the 9-tuple `*ctx` of BioFrame.iter_ctx
(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
followed by its own arguments (py.rs:195-211).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Chunk, check, lib

__all__ = ["gene_stats", "qc", "gram_accumulate", "pca_from_gram", "project", "knn_graph", "knn_query",
           "connectivities", "materialize"]


def _arr(a, dtype, name: str, writable: bool = False):
    """Typed, C-contiguous numpy view check (PyO3's typed-array extraction +
    py.rs:11-16 contiguity)."""
    if not isinstance(a, np.ndarray):
        raise TypeError(f"argument '{name}': expected a numpy array of {np.dtype(dtype)}")
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"argument '{name}': expected dtype {np.dtype(dtype)}, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name} must be C-contiguous")
    if writable and not a.flags.writeable:
        raise ValueError(f"{name} must be writable")
    return a


def _opt(a, dtype, name):
    return None if a is None else _arr(a, dtype, name)


def _p(a):
    return None if a is None or a.size == 0 else a.ctypes.data


def _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell):
    """py.rs:162-189 argument checks that need Python objects; the C side repeats the
    content checks (sortedness, indptr span)."""
    genes = _arr(genes, np.uint32, "genes")
    values = _arr(values, np.float32, "values")
    if genes.size != values.size:
        raise ValueError("genes, values must have equal length")
    cell_map = _arr(cell_map, np.int64, "cell_map")
    gene_map = _arr(gene_map, np.int32, "gene_map")
    cells = _opt(cells, np.uint32, "cells")
    indptr = _opt(indptr, np.int64, "indptr")
    if cells is not None and indptr is None and cells.size != genes.size:
        raise ValueError("cells, genes must have equal length")
    ch = Chunk(_p(genes), _p(values), genes.size, _p(cell_map), cell_map.size, _p(gene_map), gene_map.size,
               float(target_sum), int(bool(log1p)),
               None if cells is None else (cells.ctypes.data if cells.size else 1),
               None if indptr is None else (indptr.ctypes.data if indptr.size else 1),
               0 if indptr is None else indptr.size, int(first_cell))
    if indptr is not None and indptr.size == 0:
        raise ValueError("indptr must be non-decreasing and span the chunk")
    keep = (genes, values, cell_map, gene_map, cells, indptr)
    return ch, keep


def _scale_args(scale, mean, std):
    mean = _opt(mean, np.float64, "mean")
    std = _opt(std, np.float64, "std")
    return mean, std


def gene_stats(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell, out):
    """py.rs:221-232: out (n_vars, 5) += [Σx, Σx², Σexpm1 x, Σexpm1(x)², nnz]."""
    ch, _keep = _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
    out = _arr(out, np.float64, "out", writable=True)
    check(lib().crest_gene_stats(C.byref(ch), _p(out), out.size))


def qc(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell,
       flag, cell_total, cell_ngenes, cell_flag, gene_ncells, gene_total):
    """py.rs:234-248: QC accumulation on raw counts (pass target_sum=0, log1p=False):
    per row total / genes detected / total over flagged genes, per gene cells detected /
    total, all accumulated (+=)."""
    ch, _keep = _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
    flag = _arr(flag, np.bool_, "flag")
    ct = _arr(cell_total, np.float64, "cell_total", writable=True)
    cn = _arr(cell_ngenes, np.uint32, "cell_ngenes", writable=True)
    cf = _arr(cell_flag, np.float64, "cell_flag", writable=True)
    gn = _arr(gene_ncells, np.uint32, "gene_ncells", writable=True)
    gt = _arr(gene_total, np.float64, "gene_total", writable=True)
    check(lib().crest_qc(C.byref(ch), _p(flag), flag.size, _p(ct), ct.size, _p(cn), cn.size, _p(cf), cf.size,
                         _p(gn), gn.size, _p(gt), gt.size))


def gram_accumulate(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell,
                    sub_map, gram, colsum, scale, mean, std, max_value):
    """py.rs:270-287: gram (d, d) upper triangle += SᵀS, colsum += Sᵀ1 over the chunk."""
    ch, _keep = _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
    sub_map = _arr(sub_map, np.int32, "sub_map")
    gram = _arr(gram, np.float64, "gram", writable=True)
    colsum = _arr(colsum, np.float64, "colsum", writable=True)
    mean, std = _scale_args(scale, mean, std)
    check(lib().crest_gram_accumulate(
        C.byref(ch), _p(sub_map), sub_map.size, _p(gram), gram.size, _p(colsum), colsum.size, int(bool(scale)),
        _p(mean), 0 if mean is None else mean.size, _p(std), 0 if std is None else std.size,
        float(max_value or 0.0)))


def pca_from_gram(gram, colsum, n, k):
    """py.rs:386-399: (variance (k',), loadings (d, k'), total_variance)."""
    gram = _arr(gram, np.float64, "gram")
    colsum = _arr(colsum, np.float64, "colsum")
    d = colsum.size
    if n < 0 or k < 0:
        raise OverflowError("can't convert negative int to unsigned")
    kk = min(int(k), d)
    variance = np.empty(kk, np.float64)
    loadings = np.empty((d, kk), np.float64)
    total = np.zeros(1, np.float64)
    ko = C.c_size_t(0)
    check(lib().crest_pca_from_gram(_p(gram), gram.size, _p(colsum), d, int(n), int(k), _p(variance),
                                    _p(loadings), total.ctypes.data, C.byref(ko)))
    return variance, loadings, float(total[0])


def project(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell,
            sub_map, loadings, out, scale, mean, std, max_value):
    """py.rs:289-309: out[row] += (scaled row) · loadings, out (n_rows, k) float32."""
    ch, _keep = _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
    sub_map = _arr(sub_map, np.int32, "sub_map")
    loadings = _arr(loadings, np.float64, "loadings")
    if loadings.ndim != 2:
        raise ValueError("loadings must be (d, k)")
    out = _arr(out, np.float32, "out", writable=True)
    mean, std = _scale_args(scale, mean, std)
    d, k = loadings.shape
    check(lib().crest_project(
        C.byref(ch), _p(sub_map), sub_map.size, _p(loadings), d, k, _p(out), out.size, int(bool(scale)),
        _p(mean), 0 if mean is None else mean.size, _p(std), 0 if std is None else std.size,
        float(max_value or 0.0)))


def knn_graph(data, k, exact=None, seed=0):
    """py.rs:52-70: k nearest neighbours of each row, self excluded; returns
    (indices uint32 (n, k'), distances float32 (n, k')).  The GPU search is brute
    force for every n (its sets certified against the screen's error bound, see
    knn_certificate), so `exact` and `seed` do not change the result."""
    data = _arr(data, np.float32, "data")
    if data.ndim != 2:
        raise ValueError("data must be a 2-D array")
    n, dim = data.shape
    if k < 0 or seed < 0:
        raise OverflowError("can't convert negative int to unsigned")
    kk = max(min(int(k), n - 1), 1) if n > 1 else 0
    idx = np.empty((n, kk), np.uint32)
    dist = np.empty((n, kk), np.float32)
    ko = C.c_size_t(0)
    ex = -1 if exact is None else int(bool(exact))
    check(lib().crest_knn_graph(_p(data), n, dim, int(k), ex, int(seed), _p(idx), _p(dist), C.byref(ko)))
    return idx[:, :ko.value], dist[:, :ko.value]


def knn_query(reference, query, k, exact=None, seed=0):
    """py.rs:635-653: the k' = min(k, n_ref) nearest reference rows of each query row
    (knn.rs:485-546); returns (indices uint32 (n_query, k'), distances float32
    (n_query, k')) ordered by (distance, index).  The GPU search is brute force with
    certified sets for every size, so `exact` and `seed` do not change the result."""
    if not isinstance(reference, np.ndarray) or not isinstance(query, np.ndarray):
        raise TypeError("reference and query must be numpy arrays")
    if reference.ndim != 2 or query.ndim != 2:
        raise ValueError("reference and query must be 2-D arrays")
    if query.shape[1] != reference.shape[1]:
        raise ValueError("reference and query must have the same number of columns")
    reference = _arr(reference, np.float32, "reference")
    query = _arr(query, np.float32, "query")
    if k < 0 or seed < 0:
        raise OverflowError("can't convert negative int to unsigned")
    nr, dim = reference.shape
    nq = query.shape[0]
    kk = min(int(k), nr) if nq > 0 else 0
    idx = np.empty((nq, kk), np.uint32)
    dist = np.empty((nq, kk), np.float32)
    ko = C.c_size_t(0)
    ex = -1 if exact is None else int(bool(exact))
    check(lib().crest_knn_query(_p(reference), nr, _p(query), nq, dim, int(k), ex, int(seed), _p(idx), _p(dist),
                                C.byref(ko)))
    return idx[:, :ko.value], dist[:, :ko.value]


def connectivities(indices, distances, n_neighbors):
    """py.rs:72-99: UMAP fuzzy simplicial set of a kNN graph whose rows hold
    n_neighbors - 1 neighbours; returns the undirected edges (rows < cols, weights)."""
    if not isinstance(indices, np.ndarray) or not isinstance(distances, np.ndarray):
        raise TypeError("indices and distances must be numpy arrays")
    if indices.ndim != 2 or distances.shape != indices.shape:
        raise ValueError("indices and distances must be (n, k) arrays of equal shape")
    indices = _arr(indices, np.uint32, "indices")
    distances = _arr(distances, np.float32, "distances")
    if n_neighbors < 0:
        raise OverflowError("can't convert negative int to unsigned")
    n, k = indices.shape
    cap = max(n * k, 1)
    rows, cols, w = np.empty(cap, np.uint32), np.empty(cap, np.uint32), np.empty(cap, np.float32)
    m = C.c_size_t(0)
    check(lib().crest_connectivities(_p(indices), _p(distances), n, k, int(n_neighbors), _p(rows), _p(cols), _p(w),
                                     cap, C.byref(m)))
    return rows[:m.value].copy(), cols[:m.value].copy(), w[:m.value].copy()


def materialize(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell):
    """py.rs:213-219: the kept, transformed entries as (row, var, value)."""
    ch, _keep = _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell)
    n = max(int(genes.size), 1)
    rows, vars_, vals = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.float32)
    m = C.c_size_t(0)
    check(lib().crest_materialize(C.byref(ch), _p(rows), _p(vars_), _p(vals), n, C.byref(m)))
    return rows[:m.value], vars_[:m.value], vals[:m.value]
