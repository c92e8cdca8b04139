"""Fused tl.pca + pp.neighbors: one C-ABI call runs gene statistics -> mean/std ->
Gram -> FP64 eigensolve -> projection -> exact kNN on the library's stream and
reports per-stage HIP-event times.  The CSR is either resident in device memory
(DeviceCSR, crest_pipeline_run) or held in host memory and streamed through two
device slots chunk by chunk (HostCSR, crest_pipeline_run_streamed; pinned host
arrays are uploaded without a staging copy).  BatchPipeline runs many independent
cell subsets of one device CSR together (crest_pipeline_run_batched): each subset
gets exactly the result of its own Pipeline run.  Device and pinned buffers come from
the library itself (crest_dev_alloc / crest_host_alloc), so no second HIP runtime is
loaded into the process.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import PREC_NAMES, BatchArgs, BatchResult, PipelineArgs, PipelineResult, check, lib


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


class PinnedArray:
    """A numpy array in pinned host memory from crest_host_alloc, freed with the object."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        check(lib().crest_host_alloc(n, C.byref(p)))
        self.ptr = p.value
        buf = (C.c_uint8 * max(n, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                lib().crest_host_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


@dataclass
class HostCSR:
    indptr: np.ndarray   # int64 [n+1], starts at 0
    indices: np.ndarray  # uint32 [nnz]
    data: np.ndarray     # float32 [nnz]
    n_cells: int
    n_genes: int
    nnz: int
    pinned: bool
    _keep: tuple = ()

    @classmethod
    def from_arrays(cls, indptr, indices, data, n_genes: int, pinned: bool = True) -> "HostCSR":
        ip = np.asarray(indptr, np.int64)
        ip = ip - ip[0]
        ix = np.asarray(indices, np.uint32)[:int(ip[-1])]
        dv = np.asarray(data, np.float32)[:int(ip[-1])]
        if not pinned:
            arrs = [np.ascontiguousarray(a) for a in (ip, ix, dv)]
            return cls(arrs[0], arrs[1], arrs[2], len(ip) - 1, int(n_genes), int(ip[-1]), False)
        keep = []
        out = []
        for a in (ip, ix, dv):
            pa = PinnedArray(a.shape, a.dtype)
            pa.array[...] = a
            keep.append(pa)
            out.append(pa.array)
        return cls(out[0], out[1], out[2], len(ip) - 1, int(n_genes), int(ip[-1]), True, tuple(keep))


@dataclass
class PipelineOutput:
    variance: np.ndarray
    loadings: np.ndarray
    total_variance: float
    times_ms: dict


class Pipeline:
    """Reusable device outputs for repeated runs over one CSR (DeviceCSR, or HostCSR
    streamed in chunks of stream_rows cells)."""

    def __init__(self, csr, hvg_of_gene: np.ndarray, n_comps: int = 50, k: int = 15,
                 target_sum: float = 1e4, log1p: bool = True, scale: bool = True, max_value: float = 10.0,
                 prec_gram: str = "bf16x3", prec_knn: str = "bf16x3", chunk_rows: int = 65536,
                 stream_rows: int = 32768):
        self.csr = csr
        self.streamed = isinstance(csr, HostCSR)
        self.stream_rows = int(stream_rows)
        self.hvg = np.ascontiguousarray(hvg_of_gene, np.int32)
        self.d = int(self.hvg.max()) + 1
        n = csr.n_cells
        self.n = n
        self.kp = min(n_comps, self.d, n - 1)
        self.kn = max(min(k, n - 1), 1)
        self.x = DeviceBuffer(4 * n * self.kp)
        self.idx = DeviceBuffer(4 * n * self.kn)
        self.dist = DeviceBuffer(4 * n * self.kn)
        dev = (None, None, None) if self.streamed else (csr.indptr.ptr, csr.indices.ptr, csr.data.ptr)
        self.args = PipelineArgs(
            dev[0], dev[1], dev[2], n, csr.n_genes,
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
        if self.streamed:
            c = self.csr
            check(lib().crest_pipeline_run_streamed(
                C.byref(self.args), c.indptr.ctypes.data, c.indices.ctypes.data, c.data.ctypes.data,
                int(c.pinned), self.stream_rows, self.variance.ctypes.data, self.loadings.ctypes.data,
                C.byref(res)))
        else:
            check(lib().crest_pipeline_run(C.byref(self.args), self.variance.ctypes.data,
                                           self.loadings.ctypes.data, C.byref(res)))
        times = {k: getattr(res, "ms_" + k) for k in ("stats", "gram", "eig", "project", "knn_prep",
                                                       "knn_screen", "knn_rerank", "total")}
        kp = res.k_pcs
        return PipelineOutput(self.variance[:kp].copy(), self.loadings[:, :kp].copy(), res.total_variance, times)

    def x_pca(self) -> np.ndarray:
        return self.x.to_numpy(np.float32, (self.n, self.kp))

    def knn(self):
        return self.idx.to_numpy(np.uint32, (self.n, self.kn)), self.dist.to_numpy(np.float32, (self.n, self.kn))


@dataclass
class BatchOutput:
    variance: np.ndarray   # [n_subsets, k_pcs]
    loadings: np.ndarray   # [n_subsets, d, k_pcs]
    total_variance: np.ndarray  # [n_subsets]
    times_ms: dict
    tile_rows: int
    gram_rows: int
    graph_replayed: bool
    gram_tuned: bool


class BatchPipeline:
    """Independent cell subsets (cells subset_off[i] .. subset_off[i+1]-1 of a
    DeviceCSR) through gene statistics -> Gram -> eigensolve -> projection -> kNN,
    one launch per stage for all subsets.  use_graph captures the launch sequences in
    HIP graphs on the first run and replays them on later runs."""

    def __init__(self, csr: DeviceCSR, subset_off, hvg_of_gene: np.ndarray, n_comps: int = 30, k: int = 15,
                 target_sum: float = 1e4, log1p: bool = True, scale: bool = True, max_value: float = 10.0,
                 prec_gram: str = "fp16", prec_knn: str = "fp16", tile_rows: int = 0, gram_rows: int = 0,
                 eig_batch: int = 0, use_graph: bool = False):
        self.csr = csr
        self.off = np.ascontiguousarray(subset_off, np.int64)
        self.n_sub = len(self.off) - 1
        self.hvg = np.ascontiguousarray(hvg_of_gene, np.int32)
        self.d = int(self.hvg.max()) + 1
        n = csr.n_cells
        self.n = n
        self.n_comps = int(n_comps)
        self.kp = min(self.n_comps, self.d)
        self.kn = int(k)
        self.x = DeviceBuffer(4 * n * self.kp)
        self.idx = DeviceBuffer(4 * n * self.kn)
        self.dist = DeviceBuffer(4 * n * self.kn)
        pipe = PipelineArgs(
            csr.indptr.ptr, csr.indices.ptr, csr.data.ptr, n, csr.n_genes,
            self.hvg.ctypes.data, self.d, float(target_sum), int(log1p), int(scale), float(max_value or 0.0),
            self.n_comps, self.kn, PREC_NAMES[prec_gram], PREC_NAMES[prec_knn], 0,
            self.x.ptr, self.idx.ptr, self.dist.ptr)
        self.args = BatchArgs(pipe, self.n_sub, self.off.ctypes.data, int(tile_rows), int(gram_rows),
                              int(eig_batch), int(use_graph))
        self.variance = np.zeros((self.n_sub, self.n_comps), np.float64)
        self.loadings = np.zeros((self.n_sub, self.d, self.n_comps), np.float64)
        self.total = np.zeros(self.n_sub, np.float64)

    def run(self) -> BatchOutput:
        res = BatchResult()
        check(lib().crest_pipeline_run_batched(C.byref(self.args), self.variance.ctypes.data,
                                               self.loadings.ctypes.data, self.total.ctypes.data, C.byref(res)))
        times = {k: getattr(res, "ms_" + k) for k in ("stats", "gram", "eig", "project", "knn_prep",
                                                       "knn_screen", "knn_rerank", "total")}
        kp = res.k_pcs
        return BatchOutput(self.variance[:, :kp].copy(), self.loadings[:, :, :kp].copy(), self.total.copy(), times,
                           int(res.tile_rows), int(res.gram_rows), bool(res.graph_replayed), bool(res.gram_tuned))

    def x_pca(self) -> np.ndarray:
        return self.x.to_numpy(np.float32, (self.n, self.kp))

    def knn(self):
        return self.idx.to_numpy(np.uint32, (self.n, self.kn)), self.dist.to_numpy(np.float32, (self.n, self.kn))
