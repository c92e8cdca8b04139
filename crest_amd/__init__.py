"""crest_amd: CREST's streamed counts -> scale -> PCA -> kNN path on MI355X (gfx950).

The API mirrors `crest.pp.scale`, `crest.tl.pca` and `crest.pp.neighbors` of the
reference; the per-entry and dense work runs in hand-written HIP kernels of
libcrest_hip.so, reached through a C ABI (include/crest_hip.h).
"""
from . import pp, tl
from ._lib import get_knn_precision, get_precision, knn_certificate, set_knn_precision, set_precision
from .core import DEFAULT_CHUNK_NNZ, BioFrame, CSRStore, Frame, ParquetStore, TripletStore, read_parquet

__all__ = ["pp", "tl", "BioFrame", "CSRStore", "TripletStore", "ParquetStore", "read_parquet", "Frame",
           "DEFAULT_CHUNK_NNZ",
           "set_precision", "get_precision", "set_knn_precision", "get_knn_precision", "knn_certificate"]
