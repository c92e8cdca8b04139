"""Fused tl.pca + pp.neighbors over a CSR resident in device memory: one C-ABI call
(crest_pipeline_run) runs gene statistics -> mean/std -> Gram -> FP64 eigensolve ->
projection -> exact kNN on the library's stream and reports per-stage HIP-event
times.  torch provides the device buffers only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import PREC_NAMES, PipelineArgs, PipelineResult, check, lib


@dataclass
class DeviceCSR:
    indptr: "object"   # torch int64 [n+1] on the device
    indices: "object"  # torch int32 [nnz] (uint32 bit pattern)
    data: "object"     # torch float32 [nnz]
    n_cells: int
    n_genes: int

    @classmethod
    def upload(cls, indptr: np.ndarray, indices: np.ndarray, data: np.ndarray, n_genes: int, device="cuda"):
        import torch
        ip = torch.from_numpy(np.ascontiguousarray(indptr, np.int64)).to(device)
        ix = torch.from_numpy(np.ascontiguousarray(indices, np.uint32).view(np.int32)).to(device)
        dt = torch.from_numpy(np.ascontiguousarray(data, np.float32)).to(device)
        return cls(ip, ix, dt, len(indptr) - 1, int(n_genes))


@dataclass
class PipelineOutput:
    x_pca: "object"      # torch float32 [n, k_pcs]
    knn_idx: "object"    # torch int32 [n, k'] (uint32 bit pattern)
    knn_dist: "object"   # torch float32 [n, k']
    variance: np.ndarray
    loadings: np.ndarray
    total_variance: float
    times_ms: dict


class Pipeline:
    """Reusable outputs for repeated runs over one device CSR."""

    def __init__(self, csr: DeviceCSR, hvg_of_gene: np.ndarray, n_comps: int = 50, k: int = 15,
                 target_sum: float = 1e4, log1p: bool = True, scale: bool = True, max_value: float = 10.0,
                 prec_gram: str = "bf16x3", prec_knn: str = "bf16x3", chunk_rows: int = 65536):
        import torch
        self.csr = csr
        self.hvg = np.ascontiguousarray(hvg_of_gene, np.int32)
        self.d = int(self.hvg.max()) + 1
        n = csr.n_cells
        self.kp = min(n_comps, self.d, n - 1)
        self.kn = max(min(k, n - 1), 1)
        dev = csr.indptr.device
        self.x = torch.empty((n, self.kp), dtype=torch.float32, device=dev)
        self.idx = torch.empty((n, self.kn), dtype=torch.int32, device=dev)
        self.dist = torch.empty((n, self.kn), dtype=torch.float32, device=dev)
        self.args = PipelineArgs(
            csr.indptr.data_ptr(), csr.indices.data_ptr(), csr.data.data_ptr(), n, csr.n_genes,
            self.hvg.ctypes.data, self.d, float(target_sum), int(log1p), int(scale), float(max_value or 0.0),
            int(n_comps), int(k), PREC_NAMES[prec_gram], PREC_NAMES[prec_knn], int(chunk_rows),
            self.x.data_ptr(), self.idx.data_ptr(), self.dist.data_ptr())
        self.variance = np.zeros(n_comps, np.float64)
        self.loadings = np.zeros((self.d, n_comps), np.float64)

    def set_precision(self, prec_gram: str, prec_knn: str) -> None:
        self.args.prec_gram = PREC_NAMES[prec_gram]
        self.args.prec_knn = PREC_NAMES[prec_knn]

    def run(self) -> PipelineOutput:
        res = PipelineResult()
        check(lib().crest_pipeline_run(C.byref(self.args), self.variance.ctypes.data, self.loadings.ctypes.data,
                                       C.byref(res)))
        times = {k: getattr(res, "ms_" + k) for k in ("stats", "gram", "eig", "project", "knn_prep",
                                                       "knn_screen", "knn_rerank", "total")}
        kp = res.k_pcs
        return PipelineOutput(self.x, self.idx, self.dist, self.variance[:kp].copy(), self.loadings[:, :kp].copy(),
                              res.total_variance, times)
