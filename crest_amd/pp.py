"""crest.pp on the GPU library (crest/pp.py): the lazy transform record
(normalize_total, log1p, scale), QC metrics, gene statistics, highly variable genes
and the neighbour graph.

Statistics and the seurat selection run as one device session: the store's chunks
are pushed into device-resident accumulators and the selection kernel reads them
there (kernels_hvg.hip), so only the per-gene result columns come back.  Signatures,
result columns and error messages are the reference's.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _native
from .core import BioFrame, Frame
from .session import Session

__all__ = ["calculate_qc_metrics", "normalize_total", "log1p", "scale", "gene_stats", "highly_variable_genes",
           "neighbors", "connectivities_matrix"]

# cell_ranger's normalisation: the MAD of a normal sample scaled to its standard deviation
_MAD_TO_SD = 0.6744897501960817


# ----------------------------------------------------------------------------- QC
def calculate_qc_metrics(bf: BioFrame, mito_prefix: Optional[str] = "MT-", inplace: bool = True):
    """Per-cell total_counts, n_genes_by_counts, pct_counts_mt and per-gene
    n_cells_by_counts, total_counts over the raw counts of kept cells and genes
    (crest/pp.py:25-47; accumulated by the qc kernel, exact for integer counts)."""
    prefix = (mito_prefix or "").upper()
    is_mt = np.fromiter((bool(prefix) and str(name).upper().startswith(prefix) for name in bf.var_names),
                        dtype=bool, count=bf.n_vars)
    sums = {"ct": np.zeros(bf.n_obs), "cn": np.zeros(bf.n_obs, np.uint32), "cf": np.zeros(bf.n_obs),
            "gn": np.zeros(bf.n_vars, np.uint32), "gt": np.zeros(bf.n_vars)}
    for ctx in bf.iter_ctx(transform=False):
        _native.qc(*ctx, is_mt, sums["ct"], sums["cn"], sums["cf"], sums["gn"], sums["gt"])
    pct = np.zeros(bf.n_obs)
    np.divide(sums["cf"], sums["ct"], out=pct, where=sums["ct"] > 0)
    obs = Frame({"n_genes_by_counts": sums["cn"].astype(np.int64), "total_counts": sums["ct"],
                 "pct_counts_mt": pct * 100.0})
    var = Frame({"n_cells_by_counts": sums["gn"].astype(np.int64), "total_counts": sums["gt"]})
    if not inplace:
        return obs, var
    for table, cols in ((bf.obs, obs), (bf.var, var)):
        for name in cols.columns:
            table[name] = cols[name]
    return None


# ----------------------------------------------------------------------------- transform record
def normalize_total(bf: BioFrame, target_sum: float = 1e4) -> BioFrame:
    """Each cell scaled to target_sum total counts, recorded and applied lazily by the
    kernels (crest/pp.py:86-93)."""
    if not target_sum or target_sum <= 0:
        raise ValueError("target_sum must be positive")
    names = [op[0] for op in bf.ops]
    if "log1p" in names:
        raise ValueError("normalize_total must come before log1p")
    bf.ops = [op for op in bf.ops if op[0] != "normalize_total"]
    bf.ops.append(("normalize_total", float(target_sum)))
    return bf


def log1p(bf: BioFrame) -> BioFrame:
    """log(1 + x), recorded once and applied lazily (crest/pp.py:96-100)."""
    if ("log1p",) not in [tuple(op[:1]) for op in bf.ops]:
        bf.ops.append(("log1p",))
    return bf


def scale(bf: BioFrame, max_value: Optional[float] = 10.0) -> BioFrame:
    """Zero-centre, unit-variance scaling clipped at +-max_value, recorded for tl.pca,
    which applies it exactly to the sparse entries (crest/pp.py:103-111)."""
    bf.uns["scale"] = {"max_value": float(max_value) if max_value else 0.0}
    return bf


# ----------------------------------------------------------------------------- statistics / HVG
def _stats_session(bf: BioFrame) -> Session:
    """A device session holding the gene statistics of every chunk of bf."""
    sess = Session(bf.n_vars, bf.n_obs)
    for ctx in bf.iter_ctx():
        sess.push_stats(ctx)
    return sess


def gene_stats(bf: BioFrame) -> np.ndarray:
    """(n_vars, 5): Σx, Σx², Σexpm1(x), Σexpm1(x)², nnz of the transformed matrix
    (crest/pp.py:115-120)."""
    with _stats_session(bf) as sess:
        return sess.stats()


def _moments(s1: np.ndarray, s2: np.ndarray, n: int):
    """Mean and ddof-1 variance (floored at 0) from sums (crest/pp.py:123-126)."""
    mu = s1 / n
    if n > 1:
        var = (s2 - n * mu * mu) / (n - 1)
    else:
        var = np.zeros_like(mu)
    return mu, np.maximum(var, 0.0)


def _cell_ranger_columns(st: np.ndarray, n: int, n_top, min_mean, max_mean, min_disp, max_disp) -> dict:
    """cell_ranger flavour (crest/pp.py:237-253): dispersion bins at the 10th, 15th ..
    100th percentiles of the mean, normalised by the bin median and MAD."""
    m_log, v_log = _moments(st[:, 0], st[:, 1], n)
    mu = m_log.copy()
    mu[mu == 0] = 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        disp = v_log / mu
    cuts = np.concatenate(([-np.inf], np.percentile(mu, np.arange(10, 105, 5)), [np.inf]))
    code = np.clip(np.searchsorted(cuts, mu, side="left") - 1, 0, len(cuts) - 2)
    centre = np.full(len(cuts) - 1, np.nan)
    spread = np.full(len(cuts) - 1, np.nan)
    ok = ~np.isnan(disp)
    for b in np.unique(code):
        vals = disp[ok & (code == b)]
        if vals.size:
            med = np.median(vals)
            centre[b] = med
            spread[b] = np.median(np.abs(vals - med)) / _MAD_TO_SD
    with np.errstate(divide="ignore", invalid="ignore"):
        dn = (disp - centre[code]) / spread[code]
    sd_log = np.sqrt(v_log)
    sd_log[sd_log == 0] = 1.0
    return {"means": mu, "dispersions": disp, "dispersions_norm": dn,
            "highly_variable": _select(dn, mu, n_top, min_mean, max_mean, min_disp, max_disp),
            "mean_log": m_log, "std_log": sd_log}


def _select(dn, mu, n_top, min_mean, max_mean, min_disp, max_disp) -> np.ndarray:
    """The n_top largest normalised dispersions (every tie at the cut kept), or the
    mean / dispersion window (crest/pp.py:272-279)."""
    if n_top is None:
        d0 = np.nan_to_num(dn)
        return (mu > min_mean) & (mu < max_mean) & (d0 > min_disp) & (d0 < max_disp)
    finite = dn[~np.isnan(dn)]
    k = min(int(n_top), finite.size, dn.size)
    cut = np.partition(finite, finite.size - k)[finite.size - k]
    return np.where(np.isnan(dn), -np.inf, dn) >= cut


def highly_variable_genes(bf: BioFrame, n_top_genes: Optional[int] = None, flavor: str = "seurat",
                          n_bins: int = 20, min_mean: float = 0.0125, max_mean: float = 3.0,
                          min_disp: float = 0.5, max_disp: float = np.inf, subset: bool = False,
                          batch_key: Optional[str] = None) -> BioFrame:
    """sc.pp.highly_variable_genes for flavor "seurat" (crest/pp.py:201-292; gene
    statistics and selection on the device) and "cell_ranger".  Adds var columns
    means, dispersions, dispersions_norm, highly_variable, mean_log, std_log,
    n_cells_expr and records the ops they were computed under, so tl.pca reuses
    mean_log / std_log for scaling."""
    if flavor in ("seurat_v3", "seurat_v3_paper"):
        raise NotImplementedError("seurat_v3 is outside the accelerated path")
    if batch_key is not None:
        raise NotImplementedError("batch_key is supported for flavor='seurat_v3' only")
    if flavor not in ("seurat", "cell_ranger"):
        raise ValueError('flavor must be "seurat" or "cell_ranger"')
    with _stats_session(bf) as sess:
        st = sess.stats()
        if flavor == "seurat":
            cols = sess.hvg_seurat(n_top_genes, n_bins, min_mean, max_mean, min_disp, max_disp)
        else:
            cols = _cell_ranger_columns(st, bf.n_obs, n_top_genes, min_mean, max_mean, min_disp, max_disp)
    cols["n_cells_expr"] = st[:, 4].astype(np.int64)
    for name in ("means", "dispersions", "dispersions_norm", "highly_variable", "mean_log", "std_log",
                 "n_cells_expr"):
        bf.var[name] = cols[name]
    bf.uns["hvg"] = {"flavor": flavor, "ops": list(bf.ops)}
    return bf._subset(var_mask=cols["highly_variable"]) if subset else bf


# ----------------------------------------------------------------------------- neighbours
def neighbors(bf: BioFrame, n_neighbors: int = 15, n_pcs: Optional[int] = None, use_rep: str = "X_pca",
              exact: Optional[bool] = None, random_state: int = 0) -> BioFrame:
    """Exact kNN graph of obsm[use_rep] and its UMAP fuzzy connectivities
    (crest/pp.py:296-315).  n_neighbors counts the cell itself (scanpy), so each cell
    keeps n_neighbors - 1 neighbours; the search is exact for every n."""
    if use_rep not in bf.obsm:
        raise KeyError(f"{use_rep!r} not in obsm; run crest.tl.pca first")
    rep = bf.obsm[use_rep] if n_pcs is None else bf.obsm[use_rep][:, :n_pcs]
    rep = np.ascontiguousarray(rep, dtype=np.float32)
    idx, dist = _native.knn_graph(rep, n_neighbors - 1, exact=exact, seed=random_state)
    rows, cols, weights = _native.connectivities(idx, dist, n_neighbors)
    bf.uns["neighbors"] = {"params": {"n_neighbors": n_neighbors, "n_pcs": n_pcs, "use_rep": use_rep,
                                      "method": "umap"},
                           "indices": idx, "distances": dist, "rows": rows, "cols": cols, "weights": weights}
    return bf


def connectivities_matrix(bf: BioFrame):
    """The symmetric scipy.sparse CSR connectivities (scanpy's obsp['connectivities'])."""
    import scipy.sparse as sp
    g = bf.uns["neighbors"]
    upper = sp.csr_matrix((g["weights"], (g["rows"], g["cols"])), shape=(bf.n_obs, bf.n_obs))
    return (upper + upper.T).tocsr()
