"""CPU checks of the boundary and the host logic: the C ABI library loads and exports
every function include/crest_hip.h declares; argument validation raises the
reference's exception classes before any device work; the chunk protocol and the
lazy-transform record behave as crest/core.py and crest/pp.py do."""
import re
from pathlib import Path

import numpy as np
import pytest

import crest_amd as ca
from crest_amd import _lib, _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "crest_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crest_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib._SIGS), "ctypes signatures cover exactly the header"


def test_no_device_is_an_error_not_a_fallback():
    if _lib.lib().crest_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        N.knn_graph(np.zeros((4, 2), np.float32), 2)
    with pytest.raises(RuntimeError, match="no HIP device"):
        _lib.knn_certificate()


def test_screen_bound_needs_no_device():
    """The certificate's coefficients are host arithmetic: available without a GPU,
    and an unknown precision is a ValueError."""
    import ctypes as C
    coef, kpad = np.zeros(4), C.c_int32(0)
    _lib.check(_lib.lib().crest_knn_screen_bound(3, 50, coef.ctypes.data, C.byref(kpad)))
    assert kpad.value == 64 and coef[0] > 0
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().crest_knn_screen_bound(7, 50, coef.ctypes.data, C.byref(kpad)))


def _ctx(nnz=5):
    genes = np.array([0, 1, 0, 1, 2], np.uint32)[:nnz]
    vals = np.array([2, 6, 5, 1, 3], np.float32)[:nnz]
    return genes, vals, np.array([0, 1, 2], np.int64), np.array([0, 1, 2], np.int32)


def test_validation_errors_match_py_rs():
    g, v, cm, gm = _ctx()
    out = np.zeros((3, 5))
    ip = np.array([0, 2, 3, 5], np.int64)
    with pytest.raises(ValueError, match="pass exactly one of cells / indptr"):
        N.gene_stats(g, v, cm, gm, 0.0, False, None, None, 0, out)
    with pytest.raises(ValueError, match="indptr must be non-decreasing"):
        N.gene_stats(g, v, cm, gm, 0.0, False, None, np.array([0, 3, 2, 5], np.int64), 0, out)
    with pytest.raises(ValueError, match="indptr must be non-decreasing"):
        N.gene_stats(g, v, cm, gm, 0.0, False, None, np.array([0, 2, 4], np.int64), 0, out)
    with pytest.raises(ValueError, match="chunk must be sorted by cell"):
        N.gene_stats(g, v, cm, gm, 0.0, False, np.array([0, 1, 0, 1, 2], np.uint32), None, 0, out)
    with pytest.raises(ValueError, match="equal length"):
        N.gene_stats(g, v[:4], cm, gm, 0.0, False, None, ip, 0, out)
    with pytest.raises(ValueError, match=r"out must be \(n_vars, 5\)"):
        N.gene_stats(g, v, cm, gm, 0.0, False, None, ip, 0, np.zeros((2, 5)))
    with pytest.raises(ValueError, match="C-contiguous"):
        N.gene_stats(g, v, cm, gm, 0.0, False, None, ip, 0, np.zeros((5, 3)).T)
    with pytest.raises(TypeError):
        N.gene_stats(g.astype(np.int64), v, cm, gm, 0.0, False, None, ip, 0, out)
    gram, cs = np.zeros((2, 2)), np.zeros(2)
    with pytest.raises(ValueError, match="sub_map must have length n_vars"):
        N.gram_accumulate(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1], np.int32), gram, cs,
                          False, None, None, 0.0)
    with pytest.raises(ValueError, match="sub_map must have length n_vars"):
        N.gram_accumulate(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, 2], np.int32), gram, cs,
                          False, None, None, 0.0)
    with pytest.raises(ValueError, match=r"gram must be \(d, d\)"):
        N.gram_accumulate(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, -1], np.int32),
                          np.zeros((2, 3)), cs, False, None, None, 0.0)
    with pytest.raises(ValueError, match="scale=True needs mean and std"):
        N.gram_accumulate(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, -1], np.int32), gram, cs,
                          True, None, None, 0.0)
    with pytest.raises(ValueError, match="mean/std must have length d"):
        N.gram_accumulate(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, -1], np.int32), gram, cs,
                          True, np.zeros(3), np.ones(3), 0.0)
    with pytest.raises(ValueError, match=r"out must be \(n_rows, k\)"):
        N.project(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, -1], np.int32), np.zeros((2, 4)),
                  np.zeros((2, 4), np.float32), False, None, None, 0.0)
    with pytest.raises(ValueError, match=r"loadings must be \(d, k\)"):
        N.project(g, v, cm, gm, 0.0, False, None, ip, 0, np.array([0, 1, -1], np.int32), np.zeros(2),
                  np.zeros((3, 4), np.float32), False, None, None, 0.0)
    with pytest.raises(RuntimeError, match="need at least 2 cells"):
        N.pca_from_gram(np.zeros((2, 2)), np.zeros(2), 1, 2)
    with pytest.raises(ValueError, match=r"gram must be \(d, d\)"):
        N.pca_from_gram(np.zeros((2, 3)), np.zeros(2), 5, 2)


def test_knn_trivial_sizes_need_no_device():
    i, d = N.knn_graph(np.zeros((0, 3), np.float32), 5)
    assert i.shape == (0, 0)
    i, d = N.knn_graph(np.zeros((1, 3), np.float32), 5)
    assert i.shape == (1, 0) and i.dtype == np.uint32 and d.dtype == np.float32


def test_precision_switch():
    assert ca.get_precision() == "fp64"
    ca.set_precision("bf16x3")
    assert ca.get_precision() == "bf16x3"
    ca.set_precision("fp64")
    with pytest.raises(KeyError):
        ca.set_precision("int4")


def _csr(n=50, g=30, seed=0):
    rng = np.random.default_rng(seed)
    D = (rng.random((n, g)) < 0.3) * rng.integers(1, 9, (n, g))
    D[7] = 0  # an empty cell
    indptr = np.r_[0, np.cumsum((D != 0).sum(1))].astype(np.int64)
    r, c = np.nonzero(D)
    return indptr, c.astype(np.uint32), D[r, c].astype(np.float32), D


def test_csr_chunks_cover_whole_cells():
    indptr, idx, val, D = _csr()
    st = ca.CSRStore(indptr, idx, val, D.shape[1])
    seen = []
    for genes, values, cells, ip, first in st.chunks(40):
        assert cells is None and ip[0] == indptr[first] and len(genes) == ip[-1] - ip[0]
        assert len(genes) <= 40 or len(ip) == 2
        seen.extend(range(first, first + len(ip) - 1))
    assert seen == list(range(D.shape[0]))


def test_triplet_chunks_keep_cells_whole():
    indptr, idx, val, D = _csr()
    cells = np.repeat(np.arange(D.shape[0]), np.diff(indptr)).astype(np.uint32)
    st = ca.TripletStore(cells[::-1], idx[::-1], val[::-1], D.shape[0], D.shape[1])
    tot = 0
    last = -1
    for genes, values, c, ip, first in st.chunks(33):
        assert ip is None and (np.diff(c.astype(np.int64)) >= 0).all() and c[0] > last
        last = c[-1]
        tot += len(genes)
    assert tot == len(idx)


def test_iter_ctx_and_ops_record():
    indptr, idx, val, D = _csr()
    bf = ca.BioFrame.from_csr(indptr, idx, val, D.shape[1])
    bf = bf.filter_cells(np.arange(bf.n_obs) % 3 != 0)
    ca.pp.normalize_total(bf, 1e4)
    ca.pp.log1p(bf)
    ca.pp.scale(bf, None)
    assert bf.uns["scale"] == {"max_value": 0.0}
    ctx = next(bf.iter_ctx())
    assert len(ctx) == 9 and ctx[4] == 1e4 and ctx[5] is True
    assert ctx[2].dtype == np.int64 and (ctx[2] == -1).sum() == (np.arange(50) % 3 == 0).sum()
    with pytest.raises(ValueError):
        ca.pp.normalize_total(bf, 1e4)  # after log1p
    with pytest.raises(ValueError):
        ca.pp.normalize_total(bf, 0)
    assert next(bf.iter_ctx(transform=False))[4:6] == (0.0, False)


def test_pca_argument_errors():
    indptr, idx, val, D = _csr()
    bf = ca.BioFrame.from_csr(indptr, idx, val, D.shape[1])
    bf.var["highly_variable"] = np.arange(D.shape[1]) == 0
    with pytest.raises(ValueError, match="at least 2 genes"):
        ca.tl.pca(bf)
    with pytest.raises(KeyError):
        ca.pp.neighbors(bf)


def _batch_args(n_cells, off, n_comps=30, k=15):
    """crest_batch_args with placeholder device pointers: validation fails before any
    pointer is used or a device is touched."""
    import ctypes as C
    hvg = np.arange(100, dtype=np.int32)
    fake = 256
    pipe = _lib.PipelineArgs(fake, fake, fake, n_cells, 100, hvg.ctypes.data, 100, 1e4, 1, 1, 10.0, n_comps, k,
                             _lib.PREC_FP16, _lib.PREC_FP16, 0, fake, fake, fake)
    off = np.asarray(off, np.int64)
    args = _lib.BatchArgs(pipe, len(off) - 1, off.ctypes.data, 0, 0, 0, 0)
    out = np.zeros(1 << 16)
    return args, (off, hvg, out), C


@pytest.mark.parametrize("n_cells,off,n_comps,k,match", [
    (100, [0, 20, 100], 30, 15, "subset 0 has 20 cells"),
    (100, [0, 60, 99], 30, 15, "subset_off must run from 0 to n_cells"),
    (100, [0, 60, 100], 30, 80, r"1 <= k <= 64"),
    (100, [0, 50, 100], 50, 15, "subset 0 has 50 cells; a batch needs at least max\\(n_comps, k\\) \\+ 1 = 51"),
])
def test_batch_validation_errors(n_cells, off, n_comps, k, match):
    args, keep, C = _batch_args(n_cells, off, n_comps, k)
    out = keep[2]
    res = _lib.BatchResult()
    rc = _lib.lib().crest_pipeline_run_batched(C.byref(args), out.ctypes.data, out.ctypes.data, out.ctypes.data,
                                               C.byref(res))
    with pytest.raises(ValueError, match=match):
        _lib.check(rc)


def test_batch_plan_of_subsets_is_contiguous():
    """The synthetic batch generator: subsets back to back, offsets spanning all cells."""
    from crest_amd import synth
    ip, ix, dv, off = synth.make_subsets([40, 7, 90], 50, 3, density=0.1, seed=4)
    assert list(off) == [0, 40, 47, 137]
    assert len(ip) == 138 and ip[0] == 0 and ip[-1] == len(ix) == len(dv)
    a, b, c, _ = synth.make_counts(7, 50, 3, density=0.1, seed=4 + 1000 + 1)
    np.testing.assert_array_equal(ix[ip[40]:ip[47]], b)
    np.testing.assert_array_equal(np.diff(ip[40:48]), np.diff(a))


def _pipe_args(n_cells, n_genes, hvg, d):
    fake = 256
    return _lib.PipelineArgs(fake, fake, fake, n_cells, n_genes, hvg.ctypes.data, d, 1e4, 1, 1, 10.0, 10, 15,
                             _lib.PREC_FP64, _lib.PREC_FP16, 0, fake, fake, fake)


def test_streamed_rejects_decreasing_offsets_inside_a_chunk():
    """crest_pipeline_run_streamed checks every offset on the host before any upload: an
    offset inside a chunk that decreases (or overshoots the next) is a ValueError."""
    import ctypes as C
    n, g = 10, 20
    hvg = np.arange(g, dtype=np.int32)
    args = _pipe_args(n, g, hvg, g)
    ip = np.arange(0, 11 * 3, 3, dtype=np.int64)[:n + 1].copy()
    ip[4] = 40  # overshoots, then decreases at 5; the chunk boundaries 0 and 10 look fine
    genes = np.zeros(int(ip[-1]) + 64, np.uint32)
    vals = np.ones(int(ip[-1]) + 64, np.float32)
    var, load = np.zeros(10), np.zeros(g * 10)
    res = _lib.PipelineResult()
    rc = _lib.lib().crest_pipeline_run_streamed(C.byref(args), ip.ctypes.data, genes.ctypes.data, vals.ctypes.data,
                                                0, 10, var.ctypes.data, load.ctypes.data, C.byref(res))
    with pytest.raises(ValueError, match="indptr must be non-decreasing"):
        _lib.check(rc)


@pytest.mark.parametrize("entry", ["run", "streamed", "batched"])
def test_duplicate_hvg_index_is_rejected(entry):
    """hvg_of_gene may map each HVG index to one gene only (else the last gene would
    silently win mean/std while all of them fed the same Gram column)."""
    import ctypes as C
    n, g, d = 40, 20, 19
    hvg = np.arange(g, dtype=np.int32)
    hvg[19] = 3  # genes 3 and 19 both -> HVG 3
    args = _pipe_args(n, g, hvg, d)
    var, load, tot = np.zeros(64), np.zeros(g * 64), np.zeros(64)
    if entry == "run":
        rc = _lib.lib().crest_pipeline_run(C.byref(args), var.ctypes.data, load.ctypes.data,
                                           C.byref(_lib.PipelineResult()))
    elif entry == "streamed":
        ip = np.zeros(n + 1, np.int64)
        z = np.zeros(1, np.uint32)
        rc = _lib.lib().crest_pipeline_run_streamed(C.byref(args), ip.ctypes.data, z.ctypes.data, z.ctypes.data, 0,
                                                    0, var.ctypes.data, load.ctypes.data,
                                                    C.byref(_lib.PipelineResult()))
    else:
        off = np.array([0, n], np.int64)
        b = _lib.BatchArgs(args, 1, off.ctypes.data, 0, 0, 0, 0)
        rc = _lib.lib().crest_pipeline_run_batched(C.byref(b), var.ctypes.data, load.ctypes.data, tot.ctypes.data,
                                                   C.byref(_lib.BatchResult()))
    with pytest.raises(ValueError, match="maps two genes to HVG index 3"):
        _lib.check(rc)
