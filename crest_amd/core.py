"""BioFrame: raw counts in cell-sorted chunks plus annotations, with the lazy
normalize_total / log1p record of the reference (crest/core.py).

The chunk protocol is the reference's: `CSRStore.chunks(chunk_nnz)` yields
`(genes, values, cells|None, indptr|None, first_cell)` views (crest/core.py:54-63) and
`BioFrame.iter_ctx()` yields the 9-tuple the native functions take
(crest/core.py:265-278).  Annotation tables are plain column dictionaries
(`Frame`) instead of polars DataFrames (polars is not part of this build).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Iterator, Optional, Tuple

import numpy as np

DEFAULT_CHUNK_NNZ = 1 << 24  # crest/core.py:30


class Frame:
    """A minimal column table: equal-length numpy columns by name."""

    def __init__(self, cols: Optional[Dict[str, np.ndarray]] = None, height: Optional[int] = None):
        self._c: Dict[str, np.ndarray] = {}
        self._h = height
        for k, v in (cols or {}).items():
            self[k] = v

    @property
    def height(self) -> int:
        if self._h is not None:
            return self._h
        return len(next(iter(self._c.values()))) if self._c else 0

    @property
    def columns(self):
        return list(self._c)

    def __getitem__(self, k: str) -> np.ndarray:
        return self._c[k]

    def __setitem__(self, k: str, v) -> None:
        v = np.asarray(v)
        if self._c or self._h is not None:
            if len(v) != self.height:
                raise ValueError(f"column {k!r} has length {len(v)}, table has {self.height}")
        self._c[k] = v

    def __contains__(self, k: str) -> bool:
        return k in self._c

    def filter(self, mask: np.ndarray) -> "Frame":
        mask = np.asarray(mask, bool)
        return Frame({k: v[mask] for k, v in self._c.items()}, height=int(mask.sum()))

    def copy(self) -> "Frame":
        return Frame(dict(self._c), height=self._h)


class CSRStore:
    """Raw counts in memory as CSR (cells × genes), crest/core.py:34-66: ``indptr``
    int64, ``indices`` uint32 (gene ids), ``data`` float32."""

    def __init__(self, indptr, indices, data, n_genes: int):
        self.indptr = np.asarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.n_cells = len(self.indptr) - 1
        self.n_genes = int(n_genes)

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def chunks(self, chunk_nnz: int = DEFAULT_CHUNK_NNZ):
        """Whole-cell chunks of about chunk_nnz entries (crest/core.py:54-63)."""
        n = self.n_cells
        start = 0
        while start < n:
            target = self.indptr[start] + chunk_nnz
            end = int(np.searchsorted(self.indptr, target, side="right")) - 1
            end = min(max(end, start + 1), n)
            lo, hi = int(self.indptr[start]), int(self.indptr[end])
            yield self.indices[lo:hi], self.data[lo:hi], None, self.indptr[start:end + 1], start
            start = end

    def nbytes(self) -> int:
        return self.indptr.nbytes + self.indices.nbytes + self.data.nbytes


class TripletStore:
    """Raw counts as (cell, gene, value) triplets sorted by cell; chunks carry per-entry
    cell ids (the reference's FrameStore, crest/core.py:69-105)."""

    def __init__(self, cells, genes, values, n_cells: int, n_genes: int):
        cells = np.asarray(cells, np.uint32)
        order = np.argsort(cells, kind="stable") if len(cells) > 1 and np.any(cells[1:] < cells[:-1]) else None
        if order is not None:
            cells, genes, values = cells[order], np.asarray(genes)[order], np.asarray(values)[order]
        self._cells = np.ascontiguousarray(cells, np.uint32)
        self._genes = np.ascontiguousarray(genes, np.uint32)
        self._vals = np.ascontiguousarray(values, np.float32)
        self.n_cells, self.n_genes = int(n_cells), int(n_genes)

    @property
    def nnz(self) -> int:
        return len(self._cells)

    def chunks(self, chunk_nnz: int = DEFAULT_CHUNK_NNZ):
        n = len(self._cells)
        lo = 0
        while lo < n:
            hi = min(lo + chunk_nnz, n)
            if hi < n:  # extend to the end of the current cell
                hi = int(np.searchsorted(self._cells, self._cells[hi - 1], side="right"))
            yield self._genes[lo:hi], self._vals[lo:hi], self._cells[lo:hi], None, 0
            lo = hi


class ParquetStore:
    """Raw counts on disk (out-of-core), crest/core.py:108-134: a directory of
    ``part-*.parquet`` files of (cell_id, gene_id, count) triplets for ranges of whole
    cells.  One chunk = one file (``chunk_nnz`` is ignored), read with pyarrow."""

    def __init__(self, path, n_cells: int, n_genes: int):
        from pathlib import Path
        self.path = Path(path)
        self.parts = sorted(self.path.glob("part-*.parquet"))
        if not self.parts:
            raise FileNotFoundError(f"no part-*.parquet files in {self.path}")
        self.n_cells, self.n_genes = int(n_cells), int(n_genes)

    @property
    def nnz(self) -> int:
        import pyarrow.parquet as pq
        return int(sum(pq.ParquetFile(p).metadata.num_rows for p in self.parts))

    def chunks(self, chunk_nnz: int = DEFAULT_CHUNK_NNZ):
        import pyarrow.parquet as pq
        for p in self.parts:
            t = pq.read_table(p, columns=["cell_id", "gene_id", "count"])
            cols = [np.ascontiguousarray(t.column(c).to_numpy(), dt)
                    for c, dt in (("gene_id", np.uint32), ("count", np.float32), ("cell_id", np.uint32))]
            del t
            yield cols[0], cols[1], cols[2], None, 0


def _frame_to_arrow(f: "Frame"):
    import pyarrow as pa
    return pa.table({k: pa.array(np.asarray(f[k])) for k in f.columns})


def _frame_from_arrow(t) -> "Frame":
    return Frame({c: np.asarray(t.column(c).to_numpy(zero_copy_only=False)) for c in t.column_names})


@dataclass
class BioFrame:
    """A single-cell dataset: raw counts (never modified) plus annotations."""

    store: object
    obs: Frame
    var: Frame
    obsm: dict = field(default_factory=dict)
    varm: dict = field(default_factory=dict)
    uns: dict = field(default_factory=dict)
    ops: list = field(default_factory=list)  # [("normalize_total", target), ("log1p",)]
    chunk_nnz: int = DEFAULT_CHUNK_NNZ

    # ---- construction
    @classmethod
    def from_csr(cls, indptr, indices, data, n_genes: int, obs=None, var=None) -> "BioFrame":
        return cls._with_store(CSRStore(indptr, indices, data, n_genes), obs, var)

    @classmethod
    def from_scipy(cls, X, obs=None, var=None) -> "BioFrame":
        X = X.tocsr()
        X.sort_indices()
        return cls.from_csr(X.indptr, X.indices, X.data, X.shape[1], obs, var)

    @classmethod
    def from_triplets(cls, cells, genes, values, n_cells: int, n_genes: int, obs=None, var=None) -> "BioFrame":
        return cls._with_store(TripletStore(cells, genes, values, n_cells, n_genes), obs, var)

    @classmethod
    def _with_store(cls, store, obs, var) -> "BioFrame":
        obs = Frame(obs) if isinstance(obs, dict) else obs
        var = Frame(var) if isinstance(var, dict) else var
        if obs is None:
            obs = Frame({"cell_id": np.arange(store.n_cells, dtype=np.uint32)})
        elif "cell_id" not in obs:
            obs["cell_id"] = np.arange(obs.height, dtype=np.uint32)
        if var is None:
            var = Frame({"gene_id": np.arange(store.n_genes, dtype=np.uint32)})
        elif "gene_id" not in var:
            var["gene_id"] = np.arange(var.height, dtype=np.uint32)
        if obs.height != store.n_cells or var.height != store.n_genes:
            raise ValueError(f"obs/var sizes {obs.height}/{var.height} do not match store "
                             f"{store.n_cells}/{store.n_genes}")
        return cls(store=store, obs=obs, var=var)

    # ---- shape
    @property
    def n_obs(self) -> int:
        return self.obs.height

    @property
    def n_vars(self) -> int:
        return self.var.height

    @property
    def shape(self) -> Tuple[int, int]:
        return self.n_obs, self.n_vars

    @property
    def var_names(self) -> list:
        for c in ("gene_name", "gene_symbols", "gene_ids"):
            if c in self.var:
                return [str(x) for x in self.var[c]]
        return [str(x) for x in self.var["gene_id"]]

    # ---- maps from store ids to current rows (crest/core.py:244-263)
    def _cell_map(self) -> np.ndarray:
        m = np.full(self.store.n_cells, -1, dtype=np.int64)
        m[self.obs["cell_id"].astype(np.int64)] = np.arange(self.n_obs, dtype=np.int64)
        return m

    def _gene_map(self) -> np.ndarray:
        m = np.full(self.store.n_genes, -1, dtype=np.int32)
        m[self.var["gene_id"].astype(np.int64)] = np.arange(self.n_vars, dtype=np.int32)
        return m

    def _transform(self) -> Tuple[float, bool]:
        target, log = 0.0, False
        for op in self.ops:
            if op[0] == "normalize_total":
                if log:
                    raise ValueError("normalize_total after log1p is not supported")
                target = float(op[1])
            elif op[0] == "log1p":
                log = True
        return target, log

    def iter_ctx(self, transform: bool = True) -> Iterator[tuple]:
        """The 9-tuple per chunk: (genes, values, cell_map, gene_map, target_sum, log1p,
        cells, indptr, first_cell)."""
        cmap, gmap = self._cell_map(), self._gene_map()
        target, log = self._transform() if transform else (0.0, False)
        for genes, values, cells, indptr, first in self.store.chunks(self.chunk_nnz):
            yield (np.ascontiguousarray(genes, dtype=np.uint32), np.ascontiguousarray(values, dtype=np.float32),
                   cmap, gmap, target, log,
                   None if cells is None else np.ascontiguousarray(cells, dtype=np.uint32),
                   None if indptr is None else np.ascontiguousarray(indptr, dtype=np.int64), int(first))

    # ---- subsetting (crest/core.py:286-299)
    def _subset(self, obs_mask=None, var_mask=None) -> "BioFrame":
        obs = self.obs if obs_mask is None else self.obs.filter(obs_mask)
        var = self.var if var_mask is None else self.var.filter(var_mask)
        obsm = {k: v[obs_mask] for k, v in self.obsm.items()} if obs_mask is not None else dict(self.obsm)
        varm = {k: v[var_mask] for k, v in self.varm.items()} if var_mask is not None else dict(self.varm)
        uns = dict(self.uns)
        if obs_mask is not None:
            uns.pop("neighbors", None)
        return BioFrame(self.store, obs, var, obsm, varm, uns, list(self.ops), self.chunk_nnz)

    def filter_cells(self, mask) -> "BioFrame":
        return self._subset(obs_mask=np.asarray(mask, dtype=bool))

    # ---- out-of-core (crest/core.py:368-397, 418-425)
    def write_parquet(self, path, nnz_per_part: int = 1 << 23) -> None:
        """Write the raw counts of the kept cells/genes (current ids) as a directory of
        whole-cell ``part-*.parquet`` files of about nnz_per_part entries, plus
        obs.parquet / var.parquet."""
        import pyarrow as pa
        import pyarrow.parquet as pq
        from pathlib import Path
        from . import _native
        path = Path(path)
        path.mkdir(parents=True, exist_ok=True)
        for old in path.glob("part-*.parquet"):
            old.unlink()
        buf, n_buf, part = [], 0, 0

        def flush():
            nonlocal buf, n_buf, part
            if buf:
                r = np.concatenate([b[0] for b in buf])
                g = np.concatenate([b[1] for b in buf])
                v = np.concatenate([b[2] for b in buf])
                pq.write_table(pa.table({"cell_id": r, "gene_id": g, "count": v}),
                               path / f"part-{part:05d}.parquet", compression="zstd")
                part += 1
            buf, n_buf = [], 0

        for ctx in self.iter_ctx(transform=False):
            r, g, v = _native.materialize(*ctx)
            lo = 0
            while lo < len(r):
                hi = min(lo + nnz_per_part - n_buf, len(r))
                if hi < len(r):  # end on a cell boundary
                    hi = int(np.searchsorted(r, r[hi - 1], side="right"))
                buf.append((r[lo:hi], g[lo:hi], v[lo:hi]))
                n_buf += hi - lo
                lo = hi
                if n_buf >= nnz_per_part:
                    flush()
        flush()
        obs = self.obs.copy()
        obs["cell_id"] = np.arange(self.n_obs, dtype=np.uint32)
        var = self.var.copy()
        var["gene_id"] = np.arange(self.n_vars, dtype=np.uint32)
        pq.write_table(_frame_to_arrow(obs), path / "obs.parquet")
        pq.write_table(_frame_to_arrow(var), path / "var.parquet")

    def filter_genes(self, mask) -> "BioFrame":
        return self._subset(var_mask=np.asarray(mask, dtype=bool))

    def to_dense(self, transform: bool = True) -> np.ndarray:
        """Materialise the kept, transformed matrix densely (small data only)."""
        from . import _native
        D = np.zeros((self.n_obs, self.n_vars), np.float32)
        for ctx in self.iter_ctx(transform):
            r, g, x = _native.materialize(*ctx)
            D[r, g] = x
        return D


def read_parquet(path, chunk_nnz: int = DEFAULT_CHUNK_NNZ) -> BioFrame:
    """Open a dataset written by BioFrame.write_parquet (crest/core.py:418-425)."""
    import pyarrow.parquet as pq
    from pathlib import Path
    path = Path(path)
    obs = _frame_from_arrow(pq.read_table(path / "obs.parquet"))
    var = _frame_from_arrow(pq.read_table(path / "var.parquet"))
    bf = BioFrame._with_store(ParquetStore(path, obs.height, var.height), obs, var)
    bf.chunk_nnz = chunk_nnz
    return bf
