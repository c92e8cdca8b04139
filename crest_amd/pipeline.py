"""Fused tl.pca + pp.neighbors over a CSR resident in device memory: one C-ABI call
(crest_pipeline_run) runs gene statistics -> mean/std -> Gram -> FP64 eigensolve ->
projection -> exact kNN -> on the library's stream and reports per-stage HIP-event
times.  Device buffers come from the library itself (crest_dev_alloc), so no second
HIP runtime is loaded into the process.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import PREC_NAMES, PipelineArgs, PipelineResult, check, lib


class DeviceBuffer:
    """A device allocation of the library, freed with the object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().crest_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().crest_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("buffer too small for the requested array")
        check(lib().crest_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                lib().crest_dev_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


@dataclass
class DeviceCSR:
    indptr: DeviceBuffer   # int64 [n+1]
    indices: DeviceBuffer  # uint32 [nnz]
    data: DeviceBuffer     # float32 [nnz]
    n_cells: int
    n_genes: int
    nnz: int

    @classmethod
    def upload(cls, indptr: np.ndarray, indices: np.ndarray, data: np.ndarray, n_genes: int) -> "DeviceCSR":
        ip = np.ascontiguousarray(indptr, np.int64)
        if ip[0] != 0:
            ip = ip - ip[0]
        return cls(DeviceBuffer.from_array(ip), DeviceBuffer.from_array(np.ascontiguousarray(indices, np.uint32)),
                   DeviceBuffer.from_array(np.ascontiguousarray(data, np.float32)), len(ip) - 1, int(n_genes),
                   int(ip[-1]))


@dataclass
class PipelineOutput:
    variance: np.ndarray
    loadings: np.ndarray
    total_variance: float
    times_ms: dict


class Pipeline:
    """Reusable device outputs for repeated runs over one device CSR."""

    def __init__(self, csr: DeviceCSR, hvg_of_gene: np.ndarray, n_comps: int = 50, k: int = 15,
                 target_sum: float = 1e4, log1p: bool = True, scale: bool = True, max_value: float = 10.0,
                 prec_gram: str = "bf16x3", prec_knn: str = "bf16x3", chunk_rows: int = 65536):
        self.csr = csr
        self.hvg = np.ascontiguousarray(hvg_of_gene, np.int32)
        self.d = int(self.hvg.max()) + 1
        n = csr.n_cells
        self.n = n
        self.kp = min(n_comps, self.d, n - 1)
        self.kn = max(min(k, n - 1), 1)
        self.x = DeviceBuffer(4 * n * self.kp)
        self.idx = DeviceBuffer(4 * n * self.kn)
        self.dist = DeviceBuffer(4 * n * self.kn)
        self.args = PipelineArgs(
            csr.indptr.ptr, csr.indices.ptr, csr.data.ptr, n, csr.n_genes,
            self.hvg.ctypes.data, self.d, float(target_sum), int(log1p), int(scale), float(max_value or 0.0),
            int(n_comps), int(k), PREC_NAMES[prec_gram], PREC_NAMES[prec_knn], int(chunk_rows),
            self.x.ptr, self.idx.ptr, self.dist.ptr)
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
        return PipelineOutput(self.variance[:kp].copy(), self.loadings[:, :kp].copy(), res.total_variance, times)

    def x_pca(self) -> np.ndarray:
        return self.x.to_numpy(np.float32, (self.n, self.kp))

    def knn(self):
        return self.idx.to_numpy(np.uint32, (self.n, self.kn)), self.dist.to_numpy(np.float32, (self.n, self.kn))
