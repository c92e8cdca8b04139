"""crest.tl.pca on the GPU library (crest/tl.py:24-91).

The PCA of the (optionally scaled and clipped) matrix through its d x d covariance,
without forming the dense matrix: one device session takes the store's chunks through
up to three passes -- gene statistics for the scale parameters (skipped when
highly_variable_genes cached them under the same ops), the Gram, and the projection --
with the Gram, colsum, eigensolve and scores resident on the GPU.  Arguments,
results (obsm['X_pca'], varm['PCs'], uns['pca']) and errors are the reference's.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .core import BioFrame
from .session import Session

__all__ = ["pca"]

_MAX_GRAM_GENES = 20_000  # the reference's bound on the d x d Gram (crest/tl.py:55-56)


def _hvg_mask(bf: BioFrame, use_highly_variable: Optional[bool]):
    if use_highly_variable is None:
        use_highly_variable = "highly_variable" in bf.var
    if use_highly_variable:
        return np.asarray(bf.var["highly_variable"], bool), True
    return np.ones(bf.n_vars, bool), False


def _cached_scale(bf: BioFrame, mask: np.ndarray):
    """mean_log / std_log of the selected genes when highly_variable_genes computed them
    under the current ops (crest/tl.py:26-28); otherwise None (derived on the device)."""
    if bf.uns.get("hvg", {}).get("ops") != bf.ops or "mean_log" not in bf.var:
        return None, None
    return (np.ascontiguousarray(np.asarray(bf.var["mean_log"])[mask], np.float64),
            np.ascontiguousarray(np.asarray(bf.var["std_log"])[mask], np.float64))


def pca(bf: BioFrame, n_comps: int = 50, use_highly_variable: Optional[bool] = None,
        scale: Optional[bool] = None, max_value: Optional[float] = None) -> BioFrame:
    """Exact PCA of the (scaled) matrix via its covariance, streamed chunk by chunk;
    with scaling it equals sc.pp.scale(max_value) + sc.tl.pca on the dense matrix.
    Adds obsm['X_pca'] (f32 [n_obs, k]), varm['PCs'] (f32 [n_vars, k]) and uns['pca']
    (variance, variance_ratio, params, projection)."""
    mask, used_hvg = _hvg_mask(bf, use_highly_variable)
    d = int(np.count_nonzero(mask))
    if d < 2:
        raise ValueError("need at least 2 genes for PCA")
    if d > _MAX_GRAM_GENES:
        raise ValueError(f"{d} genes: select highly variable genes first (Gram PCA is O(genes^2) memory)")
    do_scale = ("scale" in bf.uns) if scale is None else bool(scale)
    if max_value is None:
        max_value = bf.uns.get("scale", {}).get("max_value", 10.0) if do_scale else 0.0
    clip = float(max_value or 0.0)
    sub = np.full(bf.n_vars, -1, np.int32)
    sub[np.flatnonzero(mask)] = np.arange(d, dtype=np.int32)
    n = bf.n_obs
    k = min(n_comps, d, n - 1)

    with Session(bf.n_vars, n) as sess:
        mean = std = None
        if do_scale:
            mean, std = _cached_scale(bf, mask)
            if mean is None:  # statistics pass; mean / std derived on the device
                for ctx in bf.iter_ctx():
                    sess.push_stats(ctx)
        sess.begin_gram(sub, d, do_scale, mean, std, clip)
        for ctx in bf.iter_ctx():
            sess.push_gram(ctx)
        variance, loadings, total, colsum, mean_used, std_used = sess.solve(k)
        shift = (colsum / n) @ loadings
        sess.begin_project(-shift.astype(np.float32))
        for ctx in bf.iter_ctx():
            sess.push_project(ctx)
        scores = sess.x_pca()

    pcs = np.zeros((bf.n_vars, k), np.float32)
    pcs[mask] = loadings
    bf.obsm["X_pca"] = scores
    bf.varm["PCs"] = pcs
    bf.uns["pca"] = {
        "variance": variance,
        "variance_ratio": variance / total,
        "params": {"scale": do_scale, "max_value": max_value, "use_highly_variable": used_hvg},
        "projection": {"genes": [g for g, m in zip(bf.var_names, mask) if m], "loadings": loadings,
                       "mean": mean_used if do_scale else None, "std": std_used if do_scale else None,
                       "shift": shift, "ops": list(bf.ops)},
    }
    return bf
