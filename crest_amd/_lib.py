"""Loader of libcrest_hip.so (the gfx950 library) and its C ABI types.

The library is built in-tree (crest_amd/csrc/build.py); every compute call goes
through it, and a missing library is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcrest_hip.so"

CREST_OK, CREST_E_VALUE, CREST_E_RUNTIME, CREST_E_DEVICE = 0, -1, -2, -3
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP16 = 0, 1, 2, 3
PREC_NAMES = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp16": PREC_FP16}


class Chunk(C.Structure):
    """struct crest_chunk (include/crest_hip.h)."""
    _fields_ = [
        ("genes", C.c_void_p), ("values", C.c_void_p), ("nnz", C.c_size_t),
        ("cell_map", C.c_void_p), ("n_cell_map", C.c_size_t),
        ("gene_map", C.c_void_p), ("n_gene_map", C.c_size_t),
        ("target_sum", C.c_double), ("log1p", C.c_int32),
        ("cells", C.c_void_p), ("indptr", C.c_void_p), ("n_indptr", C.c_size_t),
        ("first_cell", C.c_uint32),
    ]


class PipelineArgs(C.Structure):
    """struct crest_pipeline_args."""
    _fields_ = [
        ("dev_indptr", C.c_void_p), ("dev_genes", C.c_void_p), ("dev_values", C.c_void_p),
        ("n_cells", C.c_size_t), ("n_genes", C.c_size_t),
        ("hvg_of_gene", C.c_void_p), ("d", C.c_size_t),
        ("target_sum", C.c_double), ("log1p", C.c_int32), ("scale", C.c_int32), ("max_value", C.c_double),
        ("n_comps", C.c_size_t), ("k", C.c_size_t),
        ("prec_gram", C.c_int32), ("prec_knn", C.c_int32), ("chunk_rows", C.c_size_t),
        ("dev_x_pca", C.c_void_p), ("dev_knn_idx", C.c_void_p), ("dev_knn_dist", C.c_void_p),
    ]


class PipelineResult(C.Structure):
    """struct crest_pipeline_result."""
    _fields_ = [
        ("k_pcs", C.c_size_t), ("k_nn", C.c_size_t), ("total_variance", C.c_double),
        ("ms_stats", C.c_double), ("ms_gram", C.c_double), ("ms_eig", C.c_double), ("ms_project", C.c_double),
        ("ms_knn_prep", C.c_double), ("ms_knn_screen", C.c_double), ("ms_knn_rerank", C.c_double),
        ("ms_total", C.c_double),
    ]


class BatchArgs(C.Structure):
    """struct crest_batch_args."""
    _fields_ = [
        ("pipe", PipelineArgs), ("n_subsets", C.c_size_t), ("subset_off", C.c_void_p),
        ("tile_rows", C.c_size_t), ("gram_rows", C.c_size_t), ("eig_batch", C.c_size_t), ("use_graph", C.c_int32),
    ]


class BatchResult(C.Structure):
    """struct crest_batch_result."""
    _fields_ = [
        ("k_pcs", C.c_size_t), ("k_nn", C.c_size_t),
        ("ms_stats", C.c_double), ("ms_gram", C.c_double), ("ms_eig", C.c_double), ("ms_project", C.c_double),
        ("ms_knn_prep", C.c_double), ("ms_knn_screen", C.c_double), ("ms_knn_rerank", C.c_double),
        ("ms_total", C.c_double), ("tile_rows", C.c_size_t), ("gram_rows", C.c_size_t),
        ("graph_replayed", C.c_int32), ("gram_tuned", C.c_int32),
    ]


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "crest_set_device": (C.c_int, [C.c_int32]),
    "crest_set_precision": (C.c_int, [C.c_int32]),
    "crest_get_precision": (C.c_int32, []),
    "crest_set_knn_precision": (C.c_int, [C.c_int32]),
    "crest_get_knn_precision": (C.c_int32, []),
    "crest_last_error": (C.c_char_p, []),
    "crest_device_count": (C.c_int32, []),
    "crest_release": (C.c_int, []),
    "crest_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "crest_dev_free": (C.c_int, [C.c_void_p]),
    "crest_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_synchronize": (C.c_int, []),
    "crest_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_stream_synchronize": (C.c_int, [C.c_void_p]),
    "crest_gene_stats": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_size_t]),
    "crest_qc": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                           C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "crest_gram_accumulate": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                        C.c_void_p, C.c_size_t, C.c_int32, C.c_void_p, C.c_size_t, C.c_void_p,
                                        C.c_size_t, C.c_double]),
    "crest_gram_accumulate_noalloc": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                                C.c_size_t, C.c_int32, C.c_void_p, C.c_void_p, C.c_double,
                                                C.c_void_p]),
    "crest_pca_from_gram": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t)]),
    "crest_project": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                C.c_void_p, C.c_size_t, C.c_int32, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                C.c_double]),
    "crest_knn_graph": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int32, C.c_uint64,
                                  C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t)]),
    "crest_knn_query": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                  C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t)]),
    "crest_knn_certificate": (C.c_int, [C.POINTER(C.c_int64)] * 3),
    "crest_knn_screen_bound": (C.c_int, [C.c_int32, C.c_size_t, C.c_void_p, C.POINTER(C.c_int32)]),
    "crest_connectivities": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "crest_materialize": (C.c_int, [C.POINTER(Chunk), C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
    "crest_pipeline_run": (C.c_int, [C.POINTER(PipelineArgs), C.c_void_p, C.c_void_p, C.POINTER(PipelineResult)]),
    "crest_pipeline_run_streamed": (C.c_int, [C.POINTER(PipelineArgs), C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int32, C.c_size_t, C.c_void_p, C.c_void_p,
                                              C.POINTER(PipelineResult)]),
    "crest_pipeline_run_batched": (C.c_int, [C.POINTER(BatchArgs), C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.POINTER(BatchResult)]),
    "crest_session_create": (C.c_int, [C.c_size_t, C.c_size_t, C.POINTER(C.c_void_p)]),
    "crest_session_destroy": (C.c_int, [C.c_void_p]),
    "crest_session_push_stats": (C.c_int, [C.c_void_p, C.POINTER(Chunk)]),
    "crest_session_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_session_hvg_seurat": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                           C.c_double] + [C.c_void_p] * 6),
    "crest_session_begin_gram": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_int32, C.c_void_p,
                                           C.c_void_p, C.c_double]),
    "crest_session_push_gram": (C.c_int, [C.c_void_p, C.POINTER(Chunk)]),
    "crest_session_solve": (C.c_int, [C.c_void_p, C.c_size_t] + [C.c_void_p] * 6),
    "crest_session_begin_project": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_session_push_project": (C.c_int, [C.c_void_p, C.POINTER(Chunk)]),
    "crest_session_x_pca": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "crest_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "crest_host_free": (C.c_int, [C.c_void_p]),
}


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `python crest_amd/csrc/build.py` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise the reference's exception class for a non-zero status."""
    if rc == CREST_OK:
        return
    msg = (lib().crest_last_error() or b"").decode()
    if rc == CREST_E_VALUE:
        raise ValueError(msg)
    raise RuntimeError(msg)


def set_precision(prec) -> None:
    """Arithmetic of the Gram and kNN screen: "fp64" (default), "fp32", "bf16x3" or
    "fp16" (split-FP16 Gram, which needs clipped scaled values, and FP16 screen)."""
    p = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
    check(lib().crest_set_precision(p))


def get_precision() -> str:
    p = lib().crest_get_precision()
    return {v: k for k, v in PREC_NAMES.items()}[p]


def set_knn_precision(prec) -> None:
    """Arithmetic of the kNN distance screen alone: "fp64"/"fp32" (FP32 screen),
    "bf16x3", "fp16", or None to follow set_precision."""
    p = -1 if prec is None else (PREC_NAMES[prec] if isinstance(prec, str) else int(prec))
    check(lib().crest_set_knn_precision(p))


def knn_certificate() -> dict:
    """The certificate of the last kNN search on this device: `checked` queries,
    `rechecked` (the screen's error bound could not prove their sets; screened again
    by the FP32 GEMM screen), `uncertified` (still not proven: a candidate within the
    FP32 screen's error of the k-th distance, a near tie)."""
    v = [C.c_int64(0) for _ in range(3)]
    check(lib().crest_knn_certificate(*[C.byref(t) for t in v]))
    return {"checked": v[0].value, "rechecked": v[1].value, "uncertified": v[2].value}


def get_knn_precision() -> str:
    p = lib().crest_get_knn_precision()
    return {v: k for k, v in PREC_NAMES.items()}[p]
