"""Device sessions: the accumulators of one highly_variable_genes or tl.pca run kept
in GPU memory while the chunks of a BioFrame's store are pushed through pinned,
double-buffered uploads (crest_session_* in include/crest_hip.h).

The reference keeps these accumulators in numpy arrays between per-chunk native calls
(crest/pp.py:115-120, crest/tl.py:24-91): the gene statistics, the d x d Gram and
colsum, the scores.  Here they never leave the device until a result is asked for,
and the upload of each chunk overlaps the kernels of the one before.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from ._native import _chunk


class Session:
    """One session over n_vars kept genes and n_rows kept cells; use as a context
    manager (or call close()) so its device memory is released."""

    def __init__(self, n_vars: int, n_rows: int):
        self.n_vars, self.n_rows = int(n_vars), int(n_rows)
        h = C.c_void_p()
        check(lib().crest_session_create(self.n_vars, self.n_rows, C.byref(h)))
        self._h = h
        self.k = 0
        self.d = 0

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().crest_session_destroy(self._h))
        self._h = None

    def __enter__(self) -> "Session":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _push(self, fn, ctx) -> None:
        ch, _keep = _chunk(*ctx)  # the call stages the chunk before it returns
        check(fn(self._h, C.byref(ch)))

    # ---- gene statistics and highly variable genes
    def push_stats(self, ctx) -> None:
        self._push(lib().crest_session_push_stats, ctx)

    def stats(self) -> np.ndarray:
        out = np.empty((self.n_vars, 5), np.float64)
        check(lib().crest_session_stats(self._h, out.ctypes.data, out.size))
        return out

    def hvg_seurat(self, n_top, n_bins, min_mean, max_mean, min_disp, max_disp) -> dict:
        cols = {k: np.empty(self.n_vars, np.float64) for k in ("means", "dispersions", "dispersions_norm",
                                                               "mean_log", "std_log")}
        flag = np.empty(self.n_vars, np.uint8)
        check(lib().crest_session_hvg_seurat(
            self._h, -1 if n_top is None else int(n_top), int(n_bins), float(min_mean), float(max_mean),
            float(min_disp), float(max_disp), cols["means"].ctypes.data, cols["dispersions"].ctypes.data,
            cols["dispersions_norm"].ctypes.data, flag.ctypes.data, cols["mean_log"].ctypes.data,
            cols["std_log"].ctypes.data))
        cols["highly_variable"] = flag.astype(bool)
        return cols

    # ---- the PCA passes
    def begin_gram(self, sub_map: np.ndarray, d: int, scale: bool, mean, std, max_value: float) -> None:
        sub = np.ascontiguousarray(sub_map, np.int32)
        mean = None if mean is None else np.ascontiguousarray(mean, np.float64)
        std = None if std is None else np.ascontiguousarray(std, np.float64)
        check(lib().crest_session_begin_gram(self._h, sub.ctypes.data, sub.size, int(d), int(bool(scale)),
                                             None if mean is None else mean.ctypes.data,
                                             None if std is None else std.ctypes.data, float(max_value)))
        self.d = int(d)

    def push_gram(self, ctx) -> None:
        self._push(lib().crest_session_push_gram, ctx)

    def solve(self, k: int):
        """(variance [k], loadings [d, k], total, colsum [d], mean [d], std [d])."""
        d = self.d
        var, load = np.empty(k, np.float64), np.empty((d, k), np.float64)
        total, colsum = np.empty(1, np.float64), np.empty(d, np.float64)
        mean, std = np.empty(d, np.float64), np.empty(d, np.float64)
        check(lib().crest_session_solve(self._h, int(k), var.ctypes.data, load.ctypes.data, total.ctypes.data,
                                        colsum.ctypes.data, mean.ctypes.data, std.ctypes.data))
        self.k = int(k)
        return var, load, float(total[0]), colsum, mean, std

    def begin_project(self, prefill: np.ndarray) -> None:
        pre = np.ascontiguousarray(prefill, np.float32)
        check(lib().crest_session_begin_project(self._h, pre.ctypes.data, pre.size))

    def push_project(self, ctx) -> None:
        self._push(lib().crest_session_push_project, ctx)

    def x_pca(self) -> np.ndarray:
        out = np.empty((self.n_rows, self.k), np.float32)
        check(lib().crest_session_x_pca(self._h, out.ctypes.data, out.size))
        return out
