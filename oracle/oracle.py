"""ORACLE — test infrastructure, never the product.

Python side of the oracle: ctypes bindings of oracle/liboracle.so (the C
restatement of the reference's chunk kernels and exact kNN), the FP64 numpy
restatement of `pca_from_gram` (kernels.rs:311-348), dense FP64 closed forms, an
independent pandas restatement of seurat HVG selection (the scanpy algorithm that
crest/pp.py:201-292 reproduces), and the accuracy metrics.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this.

Pinning (tests/test_oracle.py):
* materialize: the reference's hand values 25/75/100 (kernels.rs:564-578) and the
  closed form log1p(D / rowsum * 1e4) (tests/test_crest.py:77-85);
* gene_stats: dense numpy sums over the transformed matrix;
* gram_accumulate + pca_from_gram: dense FP64 Z^T Z and the SVD of the clipped,
  scaled, centred dense matrix (tests/test_crest.py:88-107, kernels.rs:580-633);
  chunking invariance; a planted-mistake check;
* project: the closed form (Z - mean Z) V;
* kNN: numpy brute force argsort (tests/test_crest.py:504-511, knn.rs:399-414);
* seurat HVG: pandas.cut / groupby, the library the reference's numpy code mimics.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


class OcChunk(C.Structure):
    _fields_ = [
        ("genes", C.c_void_p), ("values", C.c_void_p), ("nnz", C.c_size_t),
        ("cell_map", C.c_void_p), ("n_cell_map", C.c_size_t),
        ("gene_map", C.c_void_p), ("n_gene_map", C.c_size_t),
        ("target_sum", C.c_double), ("log1p", C.c_int32),
        ("cells", C.c_void_p), ("indptr", C.c_void_p), ("n_indptr", C.c_size_t),
        ("first_cell", C.c_uint32),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            from . import build as _b  # type: ignore
            _b.build()
        _LIB = C.CDLL(str(path))
        _LIB.oc_materialize.restype = C.c_size_t
        for f in ("oc_set_log1p_mode", "oc_gene_stats", "oc_gram_accumulate", "oc_cov_from_gram", "oc_project",
                  "oc_knn", "oc_pair_dist", "oc_qc"):
            getattr(_LIB, f).restype = None
        _LIB.oc_gram_accumulate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                            C.c_void_p, C.c_void_p, C.c_double, C.c_int]
        _LIB.oc_cov_from_gram.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
        _LIB.oc_project.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_double]
        _LIB.oc_knn.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                C.c_void_p, C.c_void_p, C.c_int]
        _LIB.oc_pair_dist.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        _LIB.oc_materialize.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _LIB.oc_gene_stats.argtypes = [C.c_void_p, C.c_void_p]
        _LIB.oc_qc.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        _LIB.oc_set_log1p_mode.argtypes = [C.c_int]
        _LIB.oc_topk_update.restype = None
        _LIB.oc_topk_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_int64,
                                        C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int]
        _LIB.oc_scaled_moments.restype = None
        _LIB.oc_scaled_moments.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_double]
        _LIB.oc_knn_screen_f64.restype = None
        _LIB.oc_knn_screen_f64.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                           C.c_void_p, C.c_void_p, C.c_int]
        _LIB.oc_pair_dist_f64.restype = None
        _LIB.oc_pair_dist_f64.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                          C.c_int]
        _LIB.oc_connectivities.restype = C.c_size_t
        _LIB.oc_connectivities.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


def _chunk(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr, first_cell):
    keep = [genes, values, cell_map, gene_map, cells, indptr]
    ch = OcChunk(_p(genes), _p(values), len(genes), _p(cell_map), len(cell_map), _p(gene_map), len(gene_map),
                 float(target_sum), int(bool(log1p)), _p(cells), _p(indptr), 0 if indptr is None else len(indptr),
                 int(first_cell))
    return ch, keep


def set_log1p_mode(mode: int) -> None:
    """1 (default): the C library's log1pf, as the reference's f32::ln_1p
    (kernels.rs:81); 0: the correctly rounded float of the double log1p."""
    lib().oc_set_log1p_mode(int(mode))


# ------------------------------------------------------------------ chunk kernels
def materialize(*ctx):
    ch, _k = _chunk(*ctx)
    n = len(ctx[0])
    r, g, x = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.float32)
    m = lib().oc_materialize(C.byref(ch), _p(r), _p(g), _p(x))
    return r[:m], g[:m], x[:m]


def gene_stats(*ctx_out):
    *ctx, out = ctx_out
    ch, _k = _chunk(*ctx)
    lib().oc_gene_stats(C.byref(ch), _p(out))


def qc(*args):
    """kernels.rs:162-195 (crest_oracle.c oc_qc): (*ctx, flag, cell_total, cell_ngenes,
    cell_flag, gene_ncells, gene_total), all accumulated in place."""
    ctx, (flag, ct, cn, cf, gn, gt) = args[:9], args[9:]
    ch, _k = _chunk(*ctx)
    flag = np.ascontiguousarray(flag, np.uint8)
    lib().oc_qc(C.byref(ch), _p(flag), _p(ct), _p(cn), _p(cf), _p(gn), _p(gt))


def gram_accumulate(*args, threads: int = 1):
    ctx, (sub_map, gram, colsum, scale, mean, std, max_value) = args[:9], args[9:]
    ch, _k = _chunk(*ctx)
    lib().oc_gram_accumulate(C.byref(ch), _p(sub_map), _p(gram), _p(colsum), len(colsum), int(bool(scale)),
                             _p(mean), _p(std), float(max_value), int(threads))


def project(*args):
    ctx, (sub_map, loadings, out, scale, mean, std, max_value) = args[:9], args[9:]
    ch, _k = _chunk(*ctx)
    loadings = np.ascontiguousarray(loadings, np.float64)
    lib().oc_project(C.byref(ch), _p(sub_map), _p(loadings), loadings.shape[1], _p(out), int(bool(scale)),
                     _p(mean), _p(std), float(max_value))


def scaled_moments(*args):
    """(*ctx, sub_map, s1, s2, scale, mean, std, max_value): the colsum and the Gram
    diagonal of gram_accumulate, accumulated in place (crest_oracle.c oc_scaled_moments)."""
    ctx, (sub_map, s1, s2, scale, mean, std, max_value) = args[:9], args[9:]
    ch, _k = _chunk(*ctx)
    lib().oc_scaled_moments(C.byref(ch), _p(sub_map), _p(s1), _p(s2), int(bool(scale)), _p(mean), _p(std),
                            float(max_value))


def cov_from_gram(gram, colsum, n):
    d = len(colsum)
    cov = np.empty((d, d), np.float64)
    tot = np.zeros(1, np.float64)
    lib().oc_cov_from_gram(_p(gram), _p(colsum), d, int(n), _p(cov), _p(tot))
    return cov, float(tot[0])


def pca_from_gram(gram, colsum, n, k):
    """kernels.rs:311-348 with numpy LAPACK eigh: (variance desc floored at 0,
    loadings [d, k'] with the first largest-|loading| entry positive, total)."""
    d = len(colsum)
    if n < 2 or d == 0:
        raise RuntimeError("need at least 2 cells and 1 gene")
    cov, total = cov_from_gram(np.ascontiguousarray(gram, np.float64), np.ascontiguousarray(colsum, np.float64), n)
    w, u = np.linalg.eigh(cov, UPLO="U")
    k = min(k, d)
    vals = np.maximum(w[::-1][:k], 0.0)
    vecs = u[:, ::-1][:, :k].copy()
    for c in range(k):
        col = vecs[:, c]
        i = int(np.argmax(np.abs(col)))  # first maximum
        if col[i] < 0:
            vecs[:, c] = -col
    return vals, vecs, total


def knn(x, k, q0=0, q1=None, threads=1):
    """knn.rs exact path (brute force + rerank) for queries q0..q1 against all rows."""
    x = np.ascontiguousarray(x, np.float32)
    n, p = x.shape
    q1 = n if q1 is None else q1
    kk = max(min(k, n - 1), 1) if n > 1 else 0
    idx = np.empty((q1 - q0, kk), np.uint32)
    dist = np.empty((q1 - q0, kk), np.float32)
    if kk:
        lib().oc_knn(_p(x), n, p, kk, q0, q1, _p(idx), _p(dist), int(threads))
    return idx, dist


def connectivities(idx, dist, n_neighbors):
    """umap/graph.rs:61-107 + py.rs:92-97: undirected edges (rows < cols, weights)."""
    idx = np.ascontiguousarray(idx, np.uint32)
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = idx.shape
    cap = max(n * k, 1)
    r, c, w = np.empty(cap, np.uint32), np.empty(cap, np.uint32), np.empty(cap, np.float32)
    m = lib().oc_connectivities(_p(idx), _p(dist), n, k, int(n_neighbors), _p(r), _p(c), _p(w))
    return r[:m], c[:m], w[:m]


def pair_dist(x, a, b):
    x = np.ascontiguousarray(x, np.float32)
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    out = np.empty(len(a), np.float32)
    lib().oc_pair_dist(_p(x), x.shape[1], _p(a), _p(b), len(a), _p(out))
    return out


# ------------------------------------------------------------------ dense FP64 closed forms
def dense_transform(indptr, indices, data, n_genes, target_sum=1e4, log1p=True):
    """Dense normalised (+log1p) matrix in FP64 from CSR: log1p(D / rowsum * target)."""
    n = len(indptr) - 1
    D = np.zeros((n, n_genes), np.float64)
    rows = np.repeat(np.arange(n), np.diff(indptr))
    D[rows, indices] = data
    if target_sum:
        s = D.sum(1, keepdims=True)
        D = np.divide(D, s, out=np.zeros_like(D), where=s > 0) * target_sum
    return np.log1p(D) if log1p else D


def dense_scaled(D, mean, std, max_value):
    Z = (D - mean) / std
    if max_value and max_value > 0:
        Z = np.clip(Z, -max_value, max_value)
    return Z


def knn_exact_f64(x, k, queries=None, threads=None):
    """FP64 exact kNN of the given query rows of x (f32 scores, widened to f64; all
    rows as candidates, self excluded): returns (indices [m, k], squared distances [m, k+1]) -- the extra column is the
    (k+1)-th distance for tie accounting.  Ordered by (distance, index).

    Screen (oc_knn_screen_f64): FP64 d = |q|^2 + |x|^2 - 2 q.x for every pair, the
    k + 9 smallest per query kept; then every kept candidate's distance recomputed exactly as
    sum (q_j - x_j)^2 in FP64 and re-sorted.  The expanded form is within ~1e-14
    relative of the exact one, so a true neighbour is lost only if eight other points
    lie within that of the (k+1)-th distance."""
    import os
    x32 = np.ascontiguousarray(x, np.float32)
    n = len(x32)
    q = np.arange(n, dtype=np.int64) if queries is None else np.ascontiguousarray(queries, np.int64)
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = min(k + 9, n - 1)
    best_d = np.full((len(q), m), np.inf)
    best_i = np.full((len(q), m), -1, np.int64)
    L = lib()
    L.oc_knn_screen_f64(_p(x32), n, x32.shape[1], _p(q), len(q), m, _p(best_d), _p(best_i), threads)
    # exact recompute of the kept candidates, then order by (distance, index)
    a = np.repeat(q, m)
    b = best_i.ravel()
    ex = np.empty(len(a))
    L.oc_pair_dist_f64(_p(x32), x32.shape[1], _p(a), _p(np.ascontiguousarray(b)), len(a), _p(ex), threads)
    ex = ex.reshape(len(q), m)
    o = np.lexsort((best_i, ex), axis=1)
    out_i = np.take_along_axis(best_i, o, 1)[:, :k]
    out_d = np.take_along_axis(ex, o, 1)[:, :k + 1]
    return out_i, out_d


def knn_query_exact_f64(ref, query, k):
    """FP64 exact search of each query row among the reference rows (knn.rs:485-546
    with exact distances): returns (indices [nq, k], squared distances [nq, k+1]
    -- the extra column, +inf when n_ref == k, is the (k+1)-th for tie accounting),
    ordered by (distance, index)."""
    r = np.asarray(ref, np.float64)
    qx = np.asarray(query, np.float64)
    nr = len(r)
    m = min(k + 1, nr)
    nrm = (r * r).sum(1)
    out_i = np.empty((len(qx), k), np.int64)
    out_d = np.full((len(qx), k + 1), np.inf)
    for s in range(0, len(qx), 256):
        xq = qx[s:s + 256]
        d2 = (xq * xq).sum(1)[:, None] + nrm[None, :] - 2.0 * (xq @ r.T)
        part = np.argpartition(d2, m - 1, axis=1)[:, :m] if m < nr else np.tile(np.arange(nr), (len(xq), 1))
        diff = xq[:, None, :] - r[part]
        pd = (diff * diff).sum(-1)
        o = np.lexsort((part, pd), axis=1)
        part = np.take_along_axis(part, o, 1)
        pd = np.take_along_axis(pd, o, 1)
        out_i[s:s + 256] = part[:, :k]
        out_d[s:s + 256, :m] = pd
    return out_i, out_d


def knn_query_mismatch_ties(idx_got, ref, query, ref_idx, ref_d2, rtol=1e-5):
    """knn_mismatch_ties for a separate query set."""
    r = np.asarray(ref, np.float64)
    qx = np.asarray(query, np.float64)
    k = ref_idx.shape[1]
    bad = np.nonzero(~(np.sort(idx_got, 1) == np.sort(ref_idx, 1)).all(1))[0]
    ties = 0
    for b in bad:
        d = ((r[idx_got[b].astype(np.int64)] - qx[b]) ** 2).sum(1)
        if np.all(d <= ref_d2[b, k - 1] * (1 + rtol) + 1e-12):
            ties += 1
    return len(bad), ties


# ------------------------------------------------------------------ seurat HVG (pandas)
def hvg_seurat_pandas(mean_expm1, var_expm1, n_top_genes=None, n_bins=20, min_mean=0.0125, max_mean=3.0,
                      min_disp=0.5, max_disp=np.inf):
    """scanpy's seurat flavour written with pandas.cut / groupby (independent of the
    numpy restatement in the product).  Returns (highly_variable, dispersions_norm)."""
    import pandas as pd
    mean = np.asarray(mean_expm1, np.float64).copy()
    var = np.asarray(var_expm1, np.float64)
    mean[mean == 0] = 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        disp = var / mean
    disp[disp == 0] = np.nan
    with np.errstate(divide="ignore"):
        disp = np.log(disp)
    mean = np.log1p(mean)
    df = pd.DataFrame({"means": mean, "dispersions": disp})
    df["mean_bin"] = pd.cut(df["means"], bins=n_bins)
    grp = df.groupby("mean_bin", observed=False)["dispersions"]
    mb = grp.mean()
    sb = grp.std(ddof=1)
    one = sb.isnull()
    sb[one] = mb[one]
    mb[one] = 0.0
    codes = df["mean_bin"].cat.codes.to_numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        dn = (df["dispersions"].to_numpy() - mb.to_numpy()[codes]) / sb.to_numpy()[codes]
    if n_top_genes is not None:
        x = dn[~np.isnan(dn)]
        k = min(n_top_genes, x.size, len(dn))
        cutoff = np.sort(x)[::-1][k - 1]
        hv = np.nan_to_num(dn, nan=-np.inf) >= cutoff
    else:
        d0 = np.nan_to_num(dn)
        hv = (mean > min_mean) & (mean < max_mean) & (d0 > min_disp) & (d0 < max_disp)
    return hv, dn


def hvg_cell_ranger_pandas(mean_log_unused, stats, n, n_top_genes=None, min_mean=0.0125, max_mean=3.0,
                           min_disp=0.5, max_disp=np.inf):
    """scanpy's cell_ranger flavour with pandas: the log-scale mean and variance, bins
    cut at the 10th..100th percentiles of the mean (pandas.cut), per bin the median
    and the MAD / 0.6745 of the dispersions.  Returns (highly_variable, dispersions_norm)."""
    import pandas as pd
    m = stats[:, 0] / n
    v = np.maximum((stats[:, 1] - n * m * m) / (n - 1), 0.0)
    m = m.copy()
    m[m == 0] = 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        disp = v / m
    edges = np.r_[-np.inf, np.percentile(m, np.arange(10, 105, 5)), np.inf]
    df = pd.DataFrame({"means": m, "dispersions": disp})
    df["mean_bin"] = pd.cut(df["means"], edges)
    grp = df.groupby("mean_bin", observed=False)["dispersions"]
    med = grp.median()
    mad = grp.apply(lambda x: np.median(np.abs(x.dropna() - np.median(x.dropna()))) if x.notna().any()
                    else np.nan) / 0.6744897501960817
    codes = df["mean_bin"].cat.codes.to_numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        dn = (disp - med.to_numpy()[codes]) / mad.to_numpy()[codes]
    if n_top_genes is not None:
        x = dn[~np.isnan(dn)]
        k = min(n_top_genes, x.size, len(dn))
        cutoff = np.sort(x)[::-1][k - 1]
        hv = np.nan_to_num(dn, nan=-np.inf) >= cutoff
    else:
        d0 = np.nan_to_num(dn)
        hv = (m > min_mean) & (m < max_mean) & (d0 > min_disp) & (d0 < max_disp)
    return hv, dn


# ------------------------------------------------------------------ metrics
def principal_angle(V1, V2):
    """Largest principal angle (rad) between the column spans of V1 and V2, from the
    sine form arcsin(||(I - Q1 Q1^T) Q2||_2) (accurate near 0)."""
    Q1, _ = np.linalg.qr(np.asarray(V1, np.float64))
    Q2, _ = np.linalg.qr(np.asarray(V2, np.float64))
    R = Q2 - Q1 @ (Q1.T @ Q2)
    s = np.linalg.norm(R, 2)
    return float(np.arcsin(min(1.0, s)))


def knn_set_agreement(idx_a, idx_b):
    """Fraction of rows whose neighbour index sets are identical."""
    a = np.sort(np.asarray(idx_a, np.int64), 1)
    b = np.sort(np.asarray(idx_b, np.int64), 1)
    return float((a == b).all(1).mean())


def knn_jaccard(idx_a, idx_b):
    """Mean per-row Jaccard of neighbour sets (bench/paper/accuracy.py:45-47)."""
    k = idx_a.shape[1]
    tot = 0.0
    for x, y in zip(idx_a, idx_b):
        i = len(set(x.tolist()) & set(y.tolist()))
        tot += i / (2 * k - i)
    return tot / len(idx_a)


def knn_mismatch_ties(idx_got, x, queries, ref_idx, ref_d2, rtol=1e-5):
    """Rows whose sets differ, and how many of those are distance ties: every got
    neighbour's FP64 squared distance is within rtol of the reference's k-th."""
    x = np.asarray(x, np.float64)
    k = ref_idx.shape[1]
    bad = np.nonzero(~(np.sort(idx_got, 1) == np.sort(ref_idx, 1)).all(1))[0]
    ties = 0
    for r in bad:
        q = queries[r]
        d = ((x[idx_got[r].astype(np.int64)] - x[q]) ** 2).sum(1)
        kth = ref_d2[r, k - 1]
        if np.all(d <= kth * (1 + rtol) + 1e-12):
            ties += 1
    return len(bad), ties
