"""Chunk sources on the host (crest/core.py:34-134): CSR, triplets and the Parquet
directory, without a GPU."""
import numpy as np

import crest_amd as ca
from crest_amd.core import CSRStore, ParquetStore, TripletStore


def _toy(seed=0, n=60, g=25):
    rng = np.random.default_rng(seed)
    D = rng.poisson(0.8, size=(n, g)).astype(np.float32)
    nz = D != 0
    indptr = np.concatenate([[0], np.cumsum(nz.sum(1))]).astype(np.int64)
    return D, indptr, np.nonzero(nz)[1].astype(np.uint32), D[nz]


def test_csr_chunks_whole_cells():
    D, indptr, indices, data = _toy()
    st = CSRStore(indptr, indices, data, D.shape[1])
    seen = []
    for genes, vals, cells, ip, first in st.chunks(37):
        assert cells is None and ip[0] == indptr[first]
        seen.append((first, len(ip) - 1))
        assert len(genes) == ip[-1] - ip[0]
    assert seen[0][0] == 0 and sum(c for _, c in seen) == D.shape[0]


def test_parquet_store_chunks(tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq
    D, indptr, indices, data = _toy(1)
    cells = np.repeat(np.arange(D.shape[0], dtype=np.uint32), np.diff(indptr))
    cut = int(indptr[31])  # whole cells per part
    for i, (a, b) in enumerate([(0, cut), (cut, len(cells))]):
        pq.write_table(pa.table({"cell_id": cells[a:b], "gene_id": indices[a:b], "count": data[a:b]}),
                       tmp_path / f"part-{i:05d}.parquet")
    st = ParquetStore(tmp_path, D.shape[0], D.shape[1])
    assert st.nnz == len(cells)
    parts = list(st.chunks())
    assert len(parts) == 2
    g = np.concatenate([p[0] for p in parts])
    v = np.concatenate([p[1] for p in parts])
    c = np.concatenate([p[2] for p in parts])
    np.testing.assert_array_equal(g, indices)
    np.testing.assert_array_equal(v, data)
    np.testing.assert_array_equal(c, cells)
    # the same chunks as a TripletStore of the triplets
    tri = TripletStore(cells, indices, data, D.shape[0], D.shape[1])
    t = list(tri.chunks(cut))
    np.testing.assert_array_equal(t[0][2], parts[0][2])


def test_read_parquet_frames(tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq
    pq.write_table(pa.table({"cell_id": np.arange(3, dtype=np.uint32)}), tmp_path / "obs.parquet")
    pq.write_table(pa.table({"gene_id": np.arange(4, dtype=np.uint32),
                             "gene_name": np.array(["MT-1", "A", "B", "C"])}), tmp_path / "var.parquet")
    pq.write_table(pa.table({"cell_id": np.array([0, 2], np.uint32), "gene_id": np.array([1, 3], np.uint32),
                             "count": np.array([5, 7], np.float32)}), tmp_path / "part-00000.parquet")
    bf = ca.read_parquet(tmp_path)
    assert bf.n_obs == 3 and bf.n_vars == 4 and bf.var_names[0] == "MT-1"
    assert isinstance(bf.store, ParquetStore)
