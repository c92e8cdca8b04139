"""Preprocessing on the hot path (crest/pp.py): lazy normalize_total / log1p, the
scale record, per-gene statistics, seurat / cell_ranger HVG selection and the kNN
graph.  All per-entry work runs in the gfx950 library through `_native`.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _native
from .core import BioFrame, Frame

__all__ = ["calculate_qc_metrics", "normalize_total", "log1p", "scale", "gene_stats", "highly_variable_genes", "neighbors",
           "connectivities_matrix"]


def calculate_qc_metrics(bf: BioFrame, mito_prefix: Optional[str] = "MT-", inplace: bool = True):
    """pp.py:25-47: per-cell ``total_counts``, ``n_genes_by_counts``, ``pct_counts_mt``
    and per-gene ``n_cells_by_counts``, ``total_counts`` on the raw counts of the kept
    cells and genes (the accumulation runs in the GPU qc kernel)."""
    names = bf.var_names
    flag = np.array([bool(mito_prefix) and n.upper().startswith(mito_prefix.upper()) for n in names], dtype=bool)
    ct = np.zeros(bf.n_obs, np.float64)
    cn = np.zeros(bf.n_obs, np.uint32)
    cf = np.zeros(bf.n_obs, np.float64)
    gn = np.zeros(bf.n_vars, np.uint32)
    gt = np.zeros(bf.n_vars, np.float64)
    for ctx in bf.iter_ctx(transform=False):
        _native.qc(*ctx, flag, ct, cn, cf, gn, gt)
    pct = np.divide(cf, ct, out=np.zeros_like(ct), where=ct > 0) * 100.0
    obs = Frame({"n_genes_by_counts": cn.astype(np.int64), "total_counts": ct, "pct_counts_mt": pct})
    var = Frame({"n_cells_by_counts": gn.astype(np.int64), "total_counts": gt})
    if not inplace:
        return obs, var
    for k in obs.columns:
        bf.obs[k] = obs[k]
    for k in var.columns:
        bf.var[k] = var[k]
    return None


def normalize_total(bf: BioFrame, target_sum: float = 1e4) -> BioFrame:
    """Scale each cell to target_sum total counts, applied lazily (crest/pp.py:86-93)."""
    if target_sum is None or target_sum <= 0:
        raise ValueError("target_sum must be positive")
    if any(op[0] == "log1p" for op in bf.ops):
        raise ValueError("normalize_total must come before log1p")
    bf.ops = [op for op in bf.ops if op[0] != "normalize_total"] + [("normalize_total", float(target_sum))]
    return bf


def log1p(bf: BioFrame) -> BioFrame:
    """Natural log(1 + x), applied lazily (crest/pp.py:96-100)."""
    if not any(op[0] == "log1p" for op in bf.ops):
        bf.ops.append(("log1p",))
    return bf


def scale(bf: BioFrame, max_value: Optional[float] = 10.0) -> BioFrame:
    """Record zero-centre / unit-variance scaling with clipping at ±max_value; tl.pca
    applies it exactly without densifying (crest/pp.py:103-111)."""
    bf.uns["scale"] = {"max_value": float(max_value) if max_value else 0.0}
    return bf


def gene_stats(bf: BioFrame) -> np.ndarray:
    """(n_vars, 5): Σx, Σx², Σexpm1(x), Σexpm1(x)², nnz of the transformed matrix
    (crest/pp.py:115-120)."""
    out = np.zeros((bf.n_vars, 5), np.float64)
    for ctx in bf.iter_ctx():
        _native.gene_stats(*ctx, out)
    return out


def _mean_var(s1: np.ndarray, s2: np.ndarray, n: int):
    """crest/pp.py:123-126 (ddof = 1, clamped at 0)."""
    mean = s1 / n
    var = (s2 - n * mean * mean) / (n - 1) if n > 1 else np.zeros_like(mean)
    return mean, np.maximum(var, 0.0)


def _pd_cut_codes(x: np.ndarray, n_bins: int) -> np.ndarray:
    """Bin codes of pandas.cut(x, n_bins): right-closed bins, the lowest edge moved
    down by 0.1% of the range (crest/pp.py:129-136)."""
    lo, hi = float(np.min(x)), float(np.max(x))
    if lo == hi:
        return np.zeros(len(x), np.int64)
    edges = np.linspace(lo, hi, n_bins + 1)
    edges[0] -= (hi - lo) * 0.001
    return np.clip(np.searchsorted(edges, x, side="left") - 1, 0, n_bins - 1)


def highly_variable_genes(bf: BioFrame, n_top_genes: Optional[int] = None, flavor: str = "seurat",
                          n_bins: int = 20, min_mean: float = 0.0125, max_mean: float = 3.0,
                          min_disp: float = 0.5, max_disp: float = np.inf, subset: bool = False,
                          batch_key: Optional[str] = None) -> BioFrame:
    """sc.pp.highly_variable_genes for flavor "seurat" / "cell_ranger"
    (crest/pp.py:201-292).  Adds var columns means, dispersions, dispersions_norm,
    highly_variable, mean_log, std_log, n_cells_expr and caches the log-scale
    mean/std that tl.pca uses for scaling."""
    if flavor in ("seurat_v3", "seurat_v3_paper"):
        raise NotImplementedError("seurat_v3 is outside the accelerated path")
    if batch_key is not None:
        raise NotImplementedError("batch_key is supported for flavor='seurat_v3' only")
    cols = hvg_from_stats(gene_stats(bf), bf.n_obs, n_top_genes, flavor, n_bins, min_mean, max_mean,
                          min_disp, max_disp)
    for name, col in cols.items():
        bf.var[name] = col
    bf.uns["hvg"] = {"flavor": flavor, "ops": list(bf.ops)}
    if subset:
        return bf._subset(var_mask=cols["highly_variable"])
    return bf


def hvg_from_stats(st: np.ndarray, n: int, n_top_genes: Optional[int] = None, flavor: str = "seurat",
                   n_bins: int = 20, min_mean: float = 0.0125, max_mean: float = 3.0, min_disp: float = 0.5,
                   max_disp: float = np.inf) -> dict:
    """The var columns of highly_variable_genes from the (n_vars, 5) gene statistics
    of n cells (crest/pp.py:227-287)."""
    n_vars = st.shape[0]
    mean_log, var_log = _mean_var(st[:, 0], st[:, 1], n)
    if flavor == "seurat":
        mean, var = _mean_var(st[:, 2], st[:, 3], n)
    elif flavor == "cell_ranger":
        mean, var = mean_log.copy(), var_log.copy()
    else:
        raise ValueError('flavor must be "seurat" or "cell_ranger"')
    mean = mean.copy()
    mean[mean == 0] = 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        disp = var / mean
    if flavor == "seurat":
        disp[disp == 0] = np.nan
        with np.errstate(divide="ignore"):
            disp = np.log(disp)
        mean = np.log1p(mean)
        codes = _pd_cut_codes(mean, n_bins)
        nb = n_bins
    else:
        edges = np.r_[-np.inf, np.percentile(mean, np.arange(10, 105, 5)), np.inf]
        codes = np.clip(np.searchsorted(edges, mean, side="left") - 1, 0, len(edges) - 2)
        nb = len(edges) - 1

    avg = np.full(nb, np.nan)
    dev = np.full(nb, np.nan)
    for b in range(nb):
        d = disp[(codes == b) & ~np.isnan(disp)]
        if flavor == "seurat":
            if len(d):
                avg[b] = d.mean()
            if len(d) > 1:
                dev[b] = d.std(ddof=1)
        elif len(d):
            med = np.median(d)
            avg[b] = med
            dev[b] = np.median(np.abs(d - med)) / 0.6744897501960817  # statsmodels mad
    if flavor == "seurat":  # single-gene bins: normalised dispersion set to 1
        one = np.isnan(dev)
        dev[one] = avg[one]
        avg[one] = 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        disp_norm = (disp - avg[codes]) / dev[codes]

    if n_top_genes is not None:
        x = disp_norm[~np.isnan(disp_norm)]
        k = min(n_top_genes, x.size, n_vars)
        cutoff = np.sort(x)[::-1][k - 1]
        hv = np.nan_to_num(disp_norm, nan=-np.inf) >= cutoff
    else:
        dn = np.nan_to_num(disp_norm)
        hv = (mean > min_mean) & (mean < max_mean) & (dn > min_disp) & (dn < max_disp)

    std_log = np.sqrt(var_log)
    std_log[std_log == 0] = 1.0
    return {"means": mean, "dispersions": disp, "dispersions_norm": disp_norm, "highly_variable": hv,
            "mean_log": mean_log, "std_log": std_log, "n_cells_expr": st[:, 4].astype(np.int64)}


def neighbors(bf: BioFrame, n_neighbors: int = 15, n_pcs: Optional[int] = None, use_rep: str = "X_pca",
              exact: Optional[bool] = None, random_state: int = 0) -> BioFrame:
    """kNN graph + UMAP fuzzy connectivities of obsm[use_rep] (crest/pp.py:296-315).
    n_neighbors counts the cell itself, as in scanpy, so n_neighbors - 1 neighbours
    are stored.  The search is exact for every n."""
    if use_rep not in bf.obsm:
        raise KeyError(f"{use_rep!r} not in obsm; run crest_amd.tl.pca first")
    X = bf.obsm[use_rep]
    if n_pcs is not None:
        X = X[:, :n_pcs]
    X = np.ascontiguousarray(X, dtype=np.float32)
    idx, dist = _native.knn_graph(X, n_neighbors - 1, exact=exact, seed=random_state)
    rows, cols, w = _native.connectivities(idx, dist, n_neighbors)
    bf.uns["neighbors"] = {
        "params": {"n_neighbors": n_neighbors, "n_pcs": n_pcs, "use_rep": use_rep, "method": "umap"},
        "indices": idx, "distances": dist, "rows": rows, "cols": cols, "weights": w,
    }
    return bf


def connectivities_matrix(bf: BioFrame):
    """Symmetric scipy.sparse CSR connectivities (crest/pp.py:318-324)."""
    import scipy.sparse as sp
    nb = bf.uns["neighbors"]
    n = bf.n_obs
    C = sp.coo_matrix((nb["weights"], (nb["rows"], nb["cols"])), shape=(n, n)).tocsr()
    return (C + C.T).tocsr()
