"""Seeded synthetic single-cell counts: negative-binomial CSR with cluster programs.

Per cell: a cluster ~ U{0..C-1} and a library size L ~ lognormal(lib_mean, lib_sd)
times a scale calibrated to the requested density.  Per gene: a base mean
~ Gamma(gamma_shape, 1); per cluster a marker_frac of genes is up-regulated
marker_fold times.  Counts ~ NB(mean L_c m_{cl,g} / Σ_g m_{cl,g}, size nb_size).
Parameters come from numpy.random.default_rng(seed); the per-entry draws come from
a counter-based hash in libcrest_synth.so, so the matrix depends only on the seed.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcrest_synth.so"
_lib = None


def build() -> Path:
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", str(HERE / "csrc" / "synth.c"),
           "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed on csrc/synth.c:\n" + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (__graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        L.synth_count.restype = None
        L.synth_count.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_uint64,
                                  C.c_void_p, C.c_int]
        L.synth_fill.restype = None
        L.synth_fill.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_uint64,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _expected_density(lib, prob, cluster, size, sample, genes):
    mu = lib[sample, None] * prob[cluster[sample]][:, genes]
    return float((1.0 - (size / (size + mu)) ** size).mean())


def make_counts(n_cells: int, n_genes: int, n_clusters: int, density: float | None = None, seed: int = 0,
                marker_frac: float = 0.05, marker_fold: float = 8.0, gamma_shape: float = 0.3,
                lib_mean: float = 7.6, lib_sd: float = 0.5, nb_size: float = 5.0, threads: int | None = None):
    """Return (indptr int64, indices uint32, data float32, info dict)."""
    rng = np.random.default_rng(seed)
    cluster = rng.integers(0, n_clusters, n_cells).astype(np.int32)
    lib = rng.lognormal(lib_mean, lib_sd, n_cells)
    base = rng.gamma(gamma_shape, 1.0, n_genes)
    prog = np.ones((n_clusters, n_genes))
    n_mark = max(1, int(round(marker_frac * n_genes)))
    for c in range(n_clusters):
        prog[c, rng.choice(n_genes, n_mark, replace=False)] *= marker_fold
    m = prog * base[None, :]
    prob = np.ascontiguousarray(m / m.sum(1, keepdims=True))
    factor = 1.0
    if density is not None:
        sample = rng.choice(n_cells, min(n_cells, 1024), replace=False)
        genes = rng.choice(n_genes, min(n_genes, 1024), replace=False)
        lo, hi = 1e-4, 1e4
        for _ in range(40):  # bisection on the library-size scale (density is monotone in it)
            mid = np.sqrt(lo * hi)
            if _expected_density(lib * mid, prob, cluster, nb_size, sample, genes) < density:
                lo = mid
            else:
                hi = mid
        factor = np.sqrt(lo * hi)
    lib = np.ascontiguousarray(lib * factor)
    threads = threads or min(os.cpu_count() or 1, 16)
    L = _load()
    row_nnz = np.empty(n_cells, np.int64)
    L.synth_count(n_cells, n_genes, cluster.ctypes.data, lib.ctypes.data, prob.ctypes.data, nb_size, seed,
                  row_nnz.ctypes.data, threads)
    indptr = np.zeros(n_cells + 1, np.int64)
    np.cumsum(row_nnz, out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, np.uint32)
    data = np.empty(nnz, np.float32)
    L.synth_fill(n_cells, n_genes, cluster.ctypes.data, lib.ctypes.data, prob.ctypes.data, nb_size, seed,
                 indptr.ctypes.data, indices.ctypes.data, data.ctypes.data, threads)
    info = {"density": nnz / float(n_cells * n_genes), "nnz": nnz, "lib_scale": factor, "clusters": cluster}
    return indptr, indices, data, info


def make_subsets(sizes, n_genes: int, n_clusters: int, density: float | None = None, seed: int = 0,
                 threads: int | None = None):
    """Independent subsets, each from its own generator (seed + 1000 + i), concatenated
    into one CSR.  Returns (indptr int64, indices uint32, data float32, subset_off
    int64[len(sizes)+1])."""
    ips, ixs, dvs = [], [], []
    off = [0]
    base = 0
    for i, n in enumerate(sizes):
        ip, ix, dv, _ = make_counts(int(n), n_genes, n_clusters, density=density, seed=seed + 1000 + i,
                                    threads=threads)
        ips.append(ip[:-1] + base if i + 1 < len(sizes) else ip + base)
        base += int(ip[-1])
        ixs.append(ix)
        dvs.append(dv)
        off.append(off[-1] + int(n))
    return (np.concatenate(ips).astype(np.int64), np.concatenate(ixs).astype(np.uint32),
            np.concatenate(dvs).astype(np.float32), np.asarray(off, np.int64))
