"""Streamed PCA (crest/tl.py:24-91): per-gene statistics, the Gram pass, the FP64
eigensolve and the projection pass, each chunk through the gfx950 library."""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _native
from .core import BioFrame
from .pp import _mean_var, gene_stats

__all__ = ["pca"]


def _scale_params(bf: BioFrame, mask: np.ndarray):
    """Per-gene log-scale mean/std (ddof = 1), cached by HVG selection for the same
    ops (crest/tl.py:24-34)."""
    cached = bf.uns.get("hvg", {}).get("ops") == bf.ops and "mean_log" in bf.var
    if cached:
        mean, std = bf.var["mean_log"], bf.var["std_log"]
    else:
        st = gene_stats(bf)
        mean, var = _mean_var(st[:, 0], st[:, 1], bf.n_obs)
        std = np.sqrt(var)
        std[std == 0] = 1.0
    return np.ascontiguousarray(mean[mask], np.float64), np.ascontiguousarray(std[mask], np.float64)


def pca(bf: BioFrame, n_comps: int = 50, use_highly_variable: Optional[bool] = None,
        scale: Optional[bool] = None, max_value: Optional[float] = None) -> BioFrame:
    """Exact PCA of the (optionally scaled) matrix through its covariance, streamed
    chunk by chunk; equals sc.pp.scale(max_value) + sc.tl.pca on the dense matrix
    without forming it.  Adds obsm['X_pca'], varm['PCs'] and uns['pca']."""
    if use_highly_variable is None:
        use_highly_variable = "highly_variable" in bf.var
    mask = bf.var["highly_variable"].astype(bool) if use_highly_variable else np.ones(bf.n_vars, bool)
    d = int(mask.sum())
    if d < 2:
        raise ValueError("need at least 2 genes for PCA")
    if d > 20_000:
        raise ValueError(f"{d} genes: select highly variable genes first (Gram PCA is O(genes^2) memory)")
    if scale is None:
        scale = "scale" in bf.uns
    if max_value is None:
        max_value = bf.uns.get("scale", {}).get("max_value", 10.0) if scale else 0.0
    gmap = np.full(bf.n_vars, -1, np.int32)
    gmap[mask] = np.arange(d, dtype=np.int32)
    mean = std = None
    if scale:
        mean, std = _scale_params(bf, mask)

    n = bf.n_obs
    gram = np.zeros((d, d), np.float64)
    colsum = np.zeros(d, np.float64)
    for ctx in bf.iter_ctx():
        _native.gram_accumulate(*ctx, gmap, gram, colsum, scale, mean, std, float(max_value or 0.0))
    k = min(n_comps, d, n - 1)
    variance, loadings, total = _native.pca_from_gram(gram, colsum, n, k)
    del gram

    shift = (colsum / n) @ loadings
    X = np.empty((n, k), np.float32)
    X[:] = -shift.astype(np.float32)
    for ctx in bf.iter_ctx():
        _native.project(*ctx, gmap, loadings, X, scale, mean, std, float(max_value or 0.0))

    pcs = np.zeros((bf.n_vars, k), np.float32)
    pcs[mask] = loadings
    bf.obsm["X_pca"] = X
    bf.varm["PCs"] = pcs
    bf.uns["pca"] = {"variance": variance, "variance_ratio": variance / total,
                     "params": {"scale": bool(scale), "max_value": max_value,
                                "use_highly_variable": bool(use_highly_variable)},
                     "projection": {"genes": [g for g, m in zip(bf.var_names, mask) if m], "loadings": loadings,
                                    "mean": mean, "std": std, "shift": shift, "ops": list(bf.ops)}}
    return bf
