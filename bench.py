"""Benchmark: cells/s of scale + PCA + kNN at 1,000,000 cells x 2,000 HVGs on one
MI355X against the plain FP32 GPU path, at fixed kNN agreement.

A step is one pass of the hot path over the device-resident CSR: gene statistics ->
mean/std -> Gram -> FP64 eigensolve -> projection -> exact kNN (k=15), through
crest_pipeline_run.  Inputs are resident in HBM before the timed region.  The headline
runs the FP64 Gram (every one of the 50 principal components within 1e-6 rad of the
FP64 oracle) with the FP16 distance screen and FP64 rerank; the split-BF16 and
split-FP16 configurations and the plain FP32 path are timed beside it.

Prints one JSON line.  After the timed region, the cpu_baseline leg times the
oracle (the C restatement of the reference's CPU algorithm) on a bounded sample and
checks the GPU results against FP64 references (checker use only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# the CPU baseline's OpenMP threads stay on their cores (set before libgomp starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "cells/s scale+PCA+kNN @1M×2k HVG: split-BF16 vs FP32 at fixed kNN agreement"
PEAK_BF16 = 2.5e15      # MI355X dense BF16 / FP16 MFMA, FLOP/s
PEAK_FP32 = 157.3e12    # MI355X dense FP32 MFMA, FLOP/s
PREC_LABEL = {"fp64": "fp64", "fp32": "fp32", "bf16x3": "split-bf16", "fp16": "split-fp16"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def knn_certificate() -> dict:
    """The last kNN search's certificate counts (checked / rechecked / uncertified)."""
    from crest_amd._lib import knn_certificate as cert
    return cert()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(pipe, steps: int, warmup: int):
    """W untimed steps, then K timed steps; each run ends with the library's stream
    synchronized, and the timed region is bracketed by a device synchronization."""
    from crest_amd._lib import check, lib
    for _ in range(warmup):
        pipe.run()
    check(lib().crest_synchronize())
    outs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        outs.append(pipe.run().times_ms)
    check(lib().crest_synchronize())
    wall = time.perf_counter() - t0
    stages = {k: float(np.mean([o[k] for o in outs])) for k in outs[0]}
    return wall / steps * 1e3, stages


def cpu_baseline_leg(indptr, indices, data, n_genes, gpu, args):
    """Oracle (C restatement of the reference's CPU path) on a bounded sample, plus the
    accuracy checks of the GPU results against FP64 references."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    n = len(indptr) - 1
    gmap = np.arange(n_genes, dtype=np.int32)
    cmap = np.arange(n, dtype=np.int64)

    def ctx(lo, hi):
        a, b = int(indptr[lo]), int(indptr[hi])
        return (indices[a:b], data[a:b], cmap, gmap, 1e4, True, None, indptr[lo:hi + 1].copy(), lo)

    t = {}
    # stage 1: gene statistics over every cell (serial in the reference's one part)
    t0 = time.perf_counter()
    st = np.zeros((n_genes, 5))
    step = 1 << 17
    for lo in range(0, n, step):
        O.gene_stats(*ctx(lo, min(n, lo + step)), st)
    t["stats"] = time.perf_counter() - t0
    mean = st[:, 0] / n
    var = np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0)
    sd = np.sqrt(var)
    sd[sd == 0] = 1
    # stage 2: the FP64 sparse Gram over every cell (thread-owned rows, as the reference)
    t0 = time.perf_counter()
    gram, colsum = np.zeros((n_genes, n_genes)), np.zeros(n_genes)
    for lo in range(0, n, step):
        O.gram_accumulate(*ctx(lo, min(n, lo + step)), gmap, gram, colsum, True, mean, sd, 10.0, threads=threads)
    t["gram"] = time.perf_counter() - t0
    # stage 3: eigensolve (numpy LAPACK)
    t0 = time.perf_counter()
    var_o, load_o, tot_o = O.pca_from_gram(gram, colsum, n, args.pcs)
    t["eig"] = time.perf_counter() - t0
    # stage 4: projection on a sample of cells, scaled to n
    ns = min(n, 100_000)
    X = np.zeros((ns, load_o.shape[1]), np.float32)
    t0 = time.perf_counter()
    O.project(*ctx(0, ns), gmap, load_o, X, True, mean, sd, 10.0)
    t["project"] = (time.perf_counter() - t0) * n / ns
    # stage 5: exact kNN of a query sample against all n, scaled to n queries (on the
    # GPU's own X_pca so the accuracy check below can reuse it)
    Xg = gpu["x_head"]
    q = args.cpu_knn_queries
    t0 = time.perf_counter()
    O.knn(Xg, args.k, 0, q, threads=threads)
    t["knn"] = (time.perf_counter() - t0) * n / q
    total = sum(t.values())
    base = {"value": n / total, "unit": "cells/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "pinning": (f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                        f"OMP_PLACES={os.environ.get('OMP_PLACES')}, {threads} OpenMP threads"),
            "sample": (f"C restatement of kernels.rs/knn.rs: gene_stats and FP64 Gram over all {n} cells, "
                       f"numpy eigh, projection on {ns} cells and exact kNN of {q} queries x {n} points, "
                       f"both scaled linearly to {n}"),
            "stage_s": {k: round(v, 3) for k, v in t.items()}}

    # accuracy against FP64 references (checker)
    acc = {}
    ev = np.linalg.eigvalsh(O.cov_from_gram(gram, colsum, n)[0])[::-1]
    # the leading subspace whose boundary has the largest relative eigengap (r >= 2):
    # the principal angle of a subspace is only determined up to error / gap
    rel = (ev[:-1] - ev[1:]) / ev[:-1]
    r_sig = args.lead if args.lead > 0 else int(np.argmax(rel[1:args.pcs - 1]) + 2)
    acc["pca_signal_r"] = r_sig
    acc["pca_signal_rel_gap"] = float(rel[r_sig - 1])
    for tag in gpu["tags"]:
        nq = args.acc_queries if tag == "head" else args.alt_acc_queries
        qs = np.random.default_rng(0).choice(n, nq, replace=False)
        Xp = gpu[f"x_{tag}"]
        t0 = time.perf_counter()
        ref_i, ref_d2 = O.knn_exact_f64(Xp, args.k, qs, threads=threads)
        got = gpu[f"idx_{tag}"][qs]
        bad, ties = O.knn_mismatch_ties(got, Xp, qs, ref_i, ref_d2)
        log(f"accuracy {tag}: {nq} queries against the FP64 exact search in {time.perf_counter() - t0:.1f}s, "
            f"{bad} differing sets ({ties} ties)")
        acc[f"knn_set_agreement_{tag}"] = O.knn_set_agreement(got, ref_i)
        acc[f"knn_mismatch_{tag}"] = [int(bad), int(ties)]
        acc[f"knn_queries_{tag}"] = int(nq)
        L = gpu[f"load_{tag}"]
        acc[f"pca_angle_signal_{tag}"] = O.principal_angle(L[:, :r_sig], load_o[:, :r_sig])
        acc[f"pca_angle_lead20_{tag}"] = O.principal_angle(L[:, :20], load_o[:, :20])
        acc[f"pca_angle_all_{tag}"] = O.principal_angle(L, load_o)
        acc[f"variance_ratio_err_{tag}"] = float(np.max(np.abs(gpu[f"var_{tag}"] / gpu[f"tot_{tag}"]
                                                               - var_o / tot_o)))
    acc["eig_gap_p"] = float(var_o[args.pcs - 1] - ev[args.pcs])
    acc["eig_rel_gap_20"] = float(rel[19])
    return base, acc


def stage_only(args) -> int:
    """One chunk stage alone on the 1M x 2k workload, for hardware-counter runs that
    must not include the eigensolve's dispatches: 'stats' (gene statistics) or 'gram'
    (statistics, then the Gram in the headline's Gram precision), through a device
    session over the whole CSR as one chunk, or 'project' (the operator path's
    projection onto 50 fixed orthonormal loadings)."""
    import crest_amd as ca
    from crest_amd import synth
    from crest_amd.session import Session
    ip, ix, dv, info = synth.make_counts(args.cells, args.genes, args.clusters, density=args.density, seed=args.seed)
    ctx = (ix, dv, np.arange(args.cells, dtype=np.int64), np.arange(args.genes, dtype=np.int32), 1e4, True, None, ip, 0)
    ca.set_precision(args.gram_prec)
    if args.stage_only == "project":
        # the projection of every cell onto 50 orthonormal FP64 loadings (the operator
        # path's crest_project over the whole CSR as one chunk)
        from crest_amd import _native as N
        rng = np.random.default_rng(0)
        V = np.linalg.qr(rng.normal(size=(args.genes, args.pcs)))[0]
        mean, sd = np.full(args.genes, 0.5), np.full(args.genes, 1.5)
        out = np.zeros((args.cells, args.pcs), np.float32)
        for _ in range(args.steps):
            N.project(*ctx, np.arange(args.genes, dtype=np.int32), V, out, True, mean, sd, 10.0)
    else:
        with Session(args.genes, args.cells) as sess:
            for _ in range(args.steps):
                sess.push_stats(ctx)
                if args.stage_only == "gram":
                    sess.begin_gram(np.arange(args.genes, dtype=np.int32), args.genes, True, None, None, 10.0)
                    sess.push_gram(ctx)
                sess.stats()
    if args.calibrate_copy:
        from crest_amd._lib import check, lib
        from crest_amd.pipeline import DeviceBuffer
        nb = args.calibrate_copy << 20
        a, b = DeviceBuffer(nb), DeviceBuffer(nb)
        check(lib().crest_dev_copy(b.ptr, a.ptr, nb))
        check(lib().crest_synchronize())
        log(f"copy kernel: {nb} bytes read, {nb} bytes written")
    print(json.dumps({"stage_only": args.stage_only, "n": args.cells, "nnz": info["nnz"],
                      "prec_gram": args.gram_prec}), flush=True)
    return 0


def knn_only(args) -> int:
    """The kNN stage alone (knn_graph through the C ABI) on clustered synthetic scores of
    the workload's shape, plus the optional copy calibration: a short program for
    hardware-counter runs of the screen kernel."""
    import crest_amd as ca
    from crest_amd import _native as N
    rng = np.random.default_rng(args.seed)
    n, p = args.cells, args.pcs
    # PCA-score-like: cluster centres spread over the first (clusters - 1) columns
    # (variance ~3.6 each, as the 1M x 2k workload's PCs 2-30), unit noise in all p
    centers = np.zeros((args.clusters, p))
    m = min(p, args.clusters - 1)
    centers[:, :m] = rng.normal(0.0, 1.9, (args.clusters, m))
    x = (centers[rng.integers(0, args.clusters, n)] + rng.normal(0.0, 1.0, (n, p))).astype(np.float32)
    ca.set_knn_precision(args.knn_prec)
    t0 = time.perf_counter()
    idx, dist = N.knn_graph(x, args.k)
    log(f"knn_graph {n} x {p} k={args.k} ({args.knn_prec}) in {time.perf_counter() - t0:.2f}s")
    if args.calibrate_copy:
        from crest_amd._lib import check, lib
        from crest_amd.pipeline import DeviceBuffer
        nb = args.calibrate_copy << 20
        a, b = DeviceBuffer(nb), DeviceBuffer(nb)
        check(lib().crest_dev_copy(b.ptr, a.ptr, nb))
        log(f"copy kernel: {nb} bytes read, {nb} bytes written")
    print(json.dumps({"knn_only": True, "n": n, "p": p, "k": args.k, "prec": args.knn_prec}), flush=True)
    return 0


def screen_traffic(prec: str, n: int, p: int):
    """HBM bytes per launch of the screen from the committed copy-calibrated counter
    summary (profiles/knn_screen_traffic.json), when it was taken on this kernel and
    shape; otherwise None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "knn_screen_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("prec") == prec and rec.get("n") == n and rec.get("p") == p:
        return rec.get("hbm_bytes_per_launch")
    return None


def streamed_leg(args) -> dict:
    """The out-of-core configuration: the CSR stays in pinned host memory and every
    pass streams it through two device slots (crest_pipeline_run_streamed), at the
    headline precisions; checked against the oracle (FP64 Gram PCA over every cell,
    the FP64 exact kNN of sampled cells) unless --no-cpu."""
    from crest_amd import synth
    from crest_amd.pipeline import HostCSR, Pipeline
    t0 = time.perf_counter()
    ip, ix, dv, info = synth.make_counts(args.streamed_cells, args.genes, args.clusters, density=args.density,
                                         seed=args.seed + 1)
    host = HostCSR.from_arrays(ip, ix, dv, args.genes, pinned=True)
    log(f"streamed: generated {args.streamed_cells} x {args.genes}, nnz {info['nnz']} "
        f"in {time.perf_counter() - t0:.1f}s")
    pipe = Pipeline(host, np.arange(args.genes, dtype=np.int32), n_comps=args.pcs, k=args.k,
                    prec_gram=args.gram_prec, prec_knn=args.knn_prec, chunk_rows=args.stream_rows,
                    stream_rows=args.stream_rows)
    ms, stages = timed_steps(pipe, 1, 1)
    log(f"streamed ({args.streamed_cells} cells, chunks of {args.stream_rows}): {ms:.1f} ms/step, "
        f"stages {json.dumps({k: round(v, 2) for k, v in stages.items()})}")
    out = {"n_cells": args.streamed_cells, "nnz": info["nnz"], "stream_rows": args.stream_rows,
           "prec_gram": args.gram_prec, "prec_knn": args.knn_prec, "ms_per_step": ms,
           "value": args.streamed_cells / (ms * 1e-3), "unit": "cells/s",
           "stages_ms": {k: round(v, 3) for k, v in stages.items()}, "knn_certificate": knn_certificate()}
    if args.no_cpu:
        return out
    from oracle import oracle as O
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    res = pipe.run()
    X, (idx, _) = pipe.x_pca(), pipe.knn()
    del pipe, host
    t0 = time.perf_counter()
    n, g = args.streamed_cells, args.genes
    cmap, gmap = np.arange(n, dtype=np.int64), np.arange(g, dtype=np.int32)
    st = np.zeros((g, 5))
    step = 1 << 17

    def ctx(lo, hi):
        a, b = int(ip[lo]), int(ip[hi])
        return (ix[a:b], dv[a:b], cmap, gmap, 1e4, True, None, ip[lo:hi + 1].copy(), lo)

    for lo in range(0, n, step):
        O.gene_stats(*ctx(lo, min(n, lo + step)), st)
    mean = st[:, 0] / n
    sd = np.sqrt(np.maximum((st[:, 1] - n * mean * mean) / (n - 1), 0))
    sd[sd == 0] = 1
    gram, colsum = np.zeros((g, g)), np.zeros(g)
    for lo in range(0, n, step):
        O.gram_accumulate(*ctx(lo, min(n, lo + step)), gmap, gram, colsum, True, mean, sd, 10.0, threads=threads)
    var_o, load_o, tot_o = O.pca_from_gram(gram, colsum, n, args.pcs)
    qs = np.random.default_rng(0).choice(n, args.alt_acc_queries, replace=False)
    ref_i, ref_d2 = O.knn_exact_f64(X, args.k, qs, threads=threads)
    bad, ties = O.knn_mismatch_ties(idx[qs], X, qs, ref_i, ref_d2)
    out["accuracy"] = {"pca_angle_all": O.principal_angle(res.loadings, load_o),
                       "variance_ratio_err": float(np.max(np.abs(res.variance / res.total_variance
                                                                 - var_o / tot_o))),
                       "knn_set_agreement": O.knn_set_agreement(idx[qs], ref_i), "knn_mismatch": [int(bad), int(ties)],
                       "knn_queries": len(qs)}
    log(f"streamed accuracy against the oracle ({time.perf_counter() - t0:.1f}s): {out['accuracy']}")
    return out


def batched_leg(args) -> dict:
    """The batching configuration: n_subsets independent subsets of log-uniform sizes,
    each from its own NB generator, through one crest_pipeline_run_batched call
    (replayed from its captured HIP graphs), against one crest_pipeline_run per subset
    one after another.  Both at the headline precisions; inputs resident in HBM."""
    from crest_amd import synth
    from crest_amd._lib import check, lib
    from crest_amd.pipeline import BatchPipeline, DeviceCSR, Pipeline
    rng = np.random.default_rng(args.batch_seed)
    sizes = np.exp(rng.uniform(np.log(args.batch_min), np.log(args.batch_max), args.batch_subsets)).astype(np.int64)
    t0 = time.perf_counter()
    ip, ix, dv, off = synth.make_subsets(sizes, args.genes, args.batch_clusters, density=args.batch_density,
                                         seed=args.batch_seed)
    n, nnz = int(off[-1]), int(ip[-1])
    log(f"batch: generated {args.batch_subsets} subsets, {n} cells, nnz {nnz} in {time.perf_counter() - t0:.1f}s")
    hvg = np.arange(args.genes, dtype=np.int32)
    csr = DeviceCSR.upload(ip, ix, dv, args.genes)
    kw = dict(n_comps=args.batch_pcs, k=args.k, prec_gram=args.gram_prec, prec_knn=args.knn_prec)
    bkw = dict(eig_batch=args.batch_eig)
    # one eager run for the per-stage times
    pe = BatchPipeline(csr, off, hvg, **kw, **bkw)
    oe = pe.run()
    oe = pe.run()
    log(f"batch eager: {oe.times_ms['total']:.1f} ms, stages "
        f"{json.dumps({k: round(v, 2) for k, v in oe.times_ms.items()})}, gram_rows {oe.gram_rows}")
    del pe
    pg = BatchPipeline(csr, off, hvg, use_graph=True, **kw, **bkw)
    pg.run()  # captures the graphs (and tunes, if this shape is new)
    check(lib().crest_synchronize())
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        og = pg.run()
    check(lib().crest_synchronize())
    ms_b = (time.perf_counter() - t0) * 1e3 / args.batch_steps
    assert og.graph_replayed
    log(f"batch (graph replay): {ms_b:.1f} ms/step ({og.times_ms['gram']:.1f} gram graph, "
        f"{og.times_ms['eig']:.1f} eig, {og.times_ms['project']:.1f} projection+kNN graph)")
    Xb = pg.x_pca()
    ib, _ = pg.knn()
    cert_b = knn_certificate()
    base = {"workload": (f"{args.batch_subsets} subsets of log-uniform [{args.batch_min}, {args.batch_max}] cells "
                         f"({n} cells) x {args.genes} HVGs, each its own NB generator ({args.batch_clusters} clusters, "
                         f"density {nnz / n / args.genes:.4f}), PCA {args.batch_pcs}, exact kNN k={args.k}, "
                         f"seed {args.batch_seed}"),
            "n_subsets": args.batch_subsets, "n_cells": n, "nnz": nnz, "ms_batched": ms_b,
            "value_batched": n / (ms_b * 1e-3), "unit": "cells/s",
            "graph_ms": {"stats_gram": round(og.times_ms["gram"], 2), "eig": round(og.times_ms["eig"], 2),
                         "project_knn": round(og.times_ms["project"], 2)},
            "stages_ms_eager": {k: round(v, 2) for k, v in oe.times_ms.items()},
            "tile_rows": og.tile_rows, "gram_rows": og.gram_rows, "knn_certificate": cert_b}
    if args.batch_no_seq:
        return base
    # sequential: one pipeline per subset (its own CSR upload outside the timed region)
    singles = []
    for i in range(args.batch_subsets):
        a, b = int(off[i]), int(off[i + 1])
        sip = ip[a:b + 1] - ip[a]
        c1 = DeviceCSR.upload(sip, ix[ip[a]:ip[b]], dv[ip[a]:ip[b]], args.genes)
        singles.append(Pipeline(c1, hvg, chunk_rows=65536, **kw))
    big = int(np.argmax(sizes))
    singles[big].run()  # warm: the largest subset sizes every library buffer
    check(lib().crest_synchronize())
    t0 = time.perf_counter()
    outs = [s_.run() for s_ in singles]
    check(lib().crest_synchronize())
    ms_s = (time.perf_counter() - t0) * 1e3
    log(f"sequential ({args.batch_subsets} crest_pipeline_run calls): {ms_s:.1f} ms")
    # agreement of the batch with the per-subset runs
    var_rel = max(float(np.max(np.abs(og.variance[i] - o.variance) / o.variance)) for i, o in enumerate(outs))
    agree = []
    for i in range(0, args.batch_subsets, max(1, args.batch_subsets // 16)):
        a, b = int(off[i]), int(off[i + 1])
        i1, _ = singles[i].knn()
        agree.append((np.sort(ib[a:b], 1) == np.sort(i1, 1)).all(1).mean())
    seq_stage = {k: round(float(sum(o.times_ms[k] for o in outs)), 2) for k in outs[0].times_ms}
    base.update({"ms_sequential": ms_s, "speedup_vs_sequential": ms_s / ms_b, "value_sequential": n / (ms_s * 1e-3),
                 "stages_ms_sequential_sum": seq_stage, "max_variance_rel_diff_vs_sequential": var_rel,
                 "knn_set_agreement_vs_sequential": float(np.mean(agree))})
    return base


def c2_leg(args) -> dict:
    """Configuration 2 through the reference's own operator path: 100,000 cells x
    33,538 genes (20 clusters), normalize_total 1e4, log1p, seurat HVG 2,000, scale
    max_value 10, tl.pca 50, knn_graph(X, 15), chunks of 16,384 cells from host
    memory (each chunk uploaded by its call, as the reference's Python layer passes
    numpy views), plain FP32 GPU path (FP32 MFMA Gram, FP32 screen).  The second of two
    runs is timed."""
    import crest_amd as ca
    from crest_amd import _native as N
    from crest_amd import synth
    n, G = args.c2_cells, args.c2_genes
    t0 = time.perf_counter()
    ip, ix, dv, info = synth.make_counts(n, G, 20, density=0.07, seed=1)
    log(f"c2: generated {n} x {G}, nnz {info['nnz']} (density {info['density']:.4f}) in "
        f"{time.perf_counter() - t0:.1f}s")
    chunk_nnz = int(ip[min(n, 16384)])
    prev, prev_knn = ca.get_precision(), ca.get_knn_precision()
    ca.set_precision("fp32")
    ca.set_knn_precision(None)

    def run():
        bf = ca.BioFrame.from_csr(ip, ix, dv, G)
        bf.chunk_nnz = chunk_nnz
        t = {}
        ca.pp.normalize_total(bf, 1e4)
        ca.pp.log1p(bf)
        a = time.perf_counter()
        ca.pp.highly_variable_genes(bf, n_top_genes=2000)
        ca.pp.scale(bf, max_value=10)
        b = time.perf_counter()
        ca.tl.pca(bf, n_comps=50)
        c = time.perf_counter()
        idx, dist = N.knn_graph(bf.obsm["X_pca"], 15)
        e = time.perf_counter()
        t.update({"hvg": (b - a) * 1e3, "pca": (c - b) * 1e3, "knn": (e - c) * 1e3, "total": (e - a) * 1e3})
        return bf, idx, t

    try:
        run()
        bf, idx, t = run()
    finally:
        ca.set_precision(prev)
        ca.set_knn_precision(None if prev_knn == prev else prev_knn)
    log(f"c2 (fp32, operator path): {t['total']:.1f} ms, stages {json.dumps({k: round(v, 1) for k, v in t.items()})}")
    out = {"workload": (f"{n} cells x {G} genes, NB counts density {info['density']:.4f}, 20 clusters, seed 1; "
                        "normalize_total 1e4, log1p, seurat HVG 2000, scale max_value 10, PCA 50, "
                        "knn_graph(X, 15); chunks of 16,384 cells from host memory; FP32 Gram and screen"),
           "n_cells": n, "n_genes": G, "nnz": info["nnz"], "ms": t["total"], "value": n / (t["total"] * 1e-3),
           "unit": "cells/s", "stages_ms": {k: round(v, 2) for k, v in t.items()}}
    if not args.no_cpu:
        out.update(c2_cpu_baseline(ip, ix, dv, G, chunk_nnz, bf, idx, args))
    return out


def c2_cpu_baseline(ip, ix, dv, G, chunk_nnz, bf, idx, args) -> dict:
    """The oracle (C restatement of kernels.rs/knn.rs) on configuration 2: gene
    statistics, HVG selection, FP64 Gram over all cells, numpy eigh, projection of
    20,000 cells and exact kNN of 2,048 queries (both scaled to all cells); and the
    GPU results checked against it (HVG index set, variances, subspace, kNN sets)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    n = len(ip) - 1
    cmap, gall = np.arange(n, dtype=np.int64), np.arange(G, dtype=np.int32)

    def chunks():
        lo = 0
        while lo < n:
            hi = int(np.searchsorted(ip, ip[lo] + chunk_nnz, side="right")) - 1
            hi = max(min(hi, n), lo + 1)
            a, b = int(ip[lo]), int(ip[hi])
            yield lo, hi, (ix[a:b], dv[a:b], cmap, gall, 1e4, True, None, ip[lo:hi + 1].copy(), lo)
            lo = hi

    t = {}
    t0 = time.perf_counter()
    st = np.zeros((G, 5))
    for _, _, ctx in chunks():
        O.gene_stats(*ctx, st)
    m1 = st[:, 2] / n
    v1 = np.maximum((st[:, 3] - n * m1 * m1) / (n - 1), 0.0)
    hv, _ = O.hvg_seurat_pandas(m1, v1, n_top_genes=2000)
    t["hvg"] = time.perf_counter() - t0
    d = int(hv.sum())
    gmap = np.full(G, -1, np.int32)
    gmap[hv] = np.arange(d, dtype=np.int32)
    ml = st[:, 0] / n
    sdl = np.sqrt(np.maximum((st[:, 1] - n * ml * ml) / (n - 1), 0.0))
    sdl[sdl == 0] = 1.0
    mean, sd = ml[hv].copy(), sdl[hv].copy()
    t0 = time.perf_counter()
    gram, colsum = np.zeros((d, d)), np.zeros(d)
    for _, _, ctx in chunks():
        O.gram_accumulate(*ctx, gmap, gram, colsum, True, mean, sd, 10.0, threads=threads)
    t["gram"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    var_o, load_o, tot_o = O.pca_from_gram(gram, colsum, n, 50)
    t["eig"] = time.perf_counter() - t0
    ns = min(n, 20_000)
    Xs = np.zeros((ns, 50), np.float32)
    t0 = time.perf_counter()
    for lo, hi, ctx in chunks():
        if lo >= ns:
            break
        a, b = int(ip[lo]), int(ip[min(hi, ns)])
        c2 = (ix[a:b], dv[a:b], cmap[:ns], gall, 1e4, True, None, ip[lo:min(hi, ns) + 1].copy(), lo)
        O.project(*c2, gmap, load_o, Xs, True, mean, sd, 10.0)
    t["project"] = (time.perf_counter() - t0) * n / ns
    Xg = bf.obsm["X_pca"]
    q = min(n, 2048)
    t0 = time.perf_counter()
    O.knn(Xg, 15, 0, q, threads=threads)
    t["knn"] = (time.perf_counter() - t0) * n / q
    total = sum(t.values())
    # accuracy of the GPU operator path
    hv_gpu = np.asarray(bf.var["highly_variable"], bool)
    ev = np.linalg.eigvalsh(O.cov_from_gram(gram, colsum, n)[0])[::-1]
    rel = (ev[:-1] - ev[1:]) / ev[:-1]
    r = int(np.argmax(rel[1:48]) + 2)
    load_g = bf.uns["pca"]["projection"]["loadings"]
    qs = np.random.default_rng(0).choice(n, 1000, replace=False)
    ref_i, ref_d2 = O.knn_exact_f64(Xg, 15, qs)
    bad, ties = O.knn_mismatch_ties(idx[qs], Xg, qs, ref_i, ref_d2)
    acc = {"hvg_set_identical": bool((hv_gpu == hv).all()), "hvg_count": d,
           "variance_ratio_err": float(np.max(np.abs(bf.uns["pca"]["variance_ratio"] - var_o / tot_o))),
           "pca_signal_r": r, "pca_angle_signal": float(O.principal_angle(load_g[:, :r], load_o[:, :r])),
           "knn_set_agreement": float(1.0 - bad / len(qs)), "knn_mismatch_ties": [int(bad), int(ties)],
           "queries": len(qs)}
    log(f"c2 cpu baseline {total:.1f}s ({json.dumps({k: round(v, 2) for k, v in t.items()})}), accuracy {acc}")
    return {"cpu_baseline": {"value": n / total, "unit": "cells/s", "cores": threads, "kind": "port",
                             "sample": ("C restatement of kernels.rs/knn.rs: gene statistics, HVG and FP64 Gram over "
                                        f"all {n} cells, numpy eigh, projection of {ns} cells and exact kNN of {q} "
                                        f"queries x {n} points, both scaled linearly to {n}"),
                             "stage_s": {k: round(v, 3) for k, v in t.items()}},
            "accuracy": acc}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--genes", type=int, default=2000)
    ap.add_argument("--clusters", type=int, default=30)
    ap.add_argument("--density", type=float, default=0.08)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--pcs", type=int, default=50)
    ap.add_argument("--k", type=int, default=15)
    ap.add_argument("--lead", type=int, default=0,
                    help="leading-subspace size for the principal-angle check (0: largest relative eigengap)")
    ap.add_argument("--chunk-rows", type=int, default=65536)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--cpu-knn-queries", type=int, default=512)
    ap.add_argument("--acc-queries", type=int, default=65536,
                    help="sampled queries of the headline's kNN agreement check against the FP64 exact search")
    ap.add_argument("--alt-acc-queries", type=int, default=4096,
                    help="sampled queries of the other legs' kNN agreement checks")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline / accuracy leg")
    ap.add_argument("--knn-only", action="store_true", help="run only knn_graph on synthetic scores (profiling)")
    ap.add_argument("--stage-only", choices=["stats", "gram", "project"], default=None,
                    help="run only the gene statistics, statistics + Gram, or the projection over the workload (profiling)")
    ap.add_argument("--gram-prec", default="fp64", choices=["fp64", "fp32", "bf16x3", "fp16"],
                    help="Gram of the headline leg (fp64: FP64 MFMA; fp16: split-FP16 hi+lo, three MFMA products)")
    ap.add_argument("--knn-prec", default="fp16", choices=["fp32", "bf16x3", "fp16"],
                    help="kNN screen of the headline leg")
    ap.add_argument("--alt", default="bf16x3:bf16x3,fp16:fp16",
                    help="comma-separated GRAM:SCREEN precision pairs also timed (one step each) and checked; "
                         "'none' for none (bf16x3:bf16x3 = the all-split-BF16 configuration)")
    ap.add_argument("--streamed-cells", type=int, default=2_000_000, metavar="N",
                    help="also time the out-of-core form: N x genes NB counts in pinned host memory, streamed in "
                         "chunks of --stream-rows cells (reported under 'streamed', not the headline)")
    ap.add_argument("--stream-rows", type=int, default=32768)
    ap.add_argument("--batch-subsets", type=int, default=256,
                    help="also time the batching configuration with this many subsets (0: skip)")
    ap.add_argument("--batch-min", type=int, default=200)
    ap.add_argument("--batch-max", type=int, default=50000)
    ap.add_argument("--batch-clusters", type=int, default=10)
    ap.add_argument("--batch-density", type=float, default=0.07)
    ap.add_argument("--batch-pcs", type=int, default=30)
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-eig", type=int, default=0, help="subsets per batched eigensolve call (0: library default)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batching configuration (profiling)")
    ap.add_argument("--c2-cells", type=int, default=100_000,
                    help="also run configuration 2 (this many cells x --c2-genes, HVG selection, FP32 operator "
                         "path) with its CPU reference (0: skip)")
    ap.add_argument("--c2-genes", type=int, default=33_538)
    ap.add_argument("--c2-only", action="store_true", help="run only configuration 2")
    ap.add_argument("--batch-no-seq", action="store_true", help="skip the sequential comparison of the batch leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run one plain device copy of MB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    from crest_amd import synth
    from crest_amd.pipeline import DeviceCSR, Pipeline

    if args.knn_only:
        return knn_only(args)
    if args.stage_only:
        return stage_only(args)
    if args.batch_only:
        print(json.dumps({"batch_only": True, "batch": batched_leg(args)}), flush=True)
        return 0
    if args.c2_only:
        print(json.dumps({"c2_only": True, "c2": c2_leg(args)}), flush=True)
        return 0

    t0 = time.perf_counter()
    indptr, indices, data, info = synth.make_counts(args.cells, args.genes, args.clusters, density=args.density,
                                                    seed=args.seed)
    log(f"generated {args.cells} x {args.genes}, nnz {info['nnz']} (density {info['density']:.4f}) "
        f"in {time.perf_counter() - t0:.1f}s")
    csr = DeviceCSR.upload(indptr, indices, data, args.genes)
    hvg = np.arange(args.genes, dtype=np.int32)  # the 2,000 genes are the pre-selected HVGs
    pipe = Pipeline(csr, hvg, n_comps=args.pcs, k=args.k, prec_gram=args.gram_prec, prec_knn=args.knn_prec,
                    chunk_rows=args.chunk_rows)
    gpu = {"tags": []}

    def capture(tag):
        out_ = pipe.run()
        gpu["tags"].append(tag)
        gpu.update({f"x_{tag}": pipe.x_pca(), f"idx_{tag}": pipe.knn()[0], f"var_{tag}": out_.variance,
                    f"load_{tag}": out_.loadings, f"tot_{tag}": out_.total_variance,
                    f"cert_{tag}": knn_certificate()})

    ms, stages = timed_steps(pipe, args.steps, args.warmup)
    log(f"{args.gram_prec} Gram + {args.knn_prec} screen: {ms:.1f} ms/step, "
        f"stages {json.dumps({k: round(v, 2) for k, v in stages.items()})}")
    capture("head")
    alts = []
    for pair in ([] if args.alt == "none" else args.alt.split(",")):
        pg, pk = pair.split(":")
        if (pg, pk) == (args.gram_prec, args.knn_prec):
            continue
        tag = f"{pg}_{pk}"
        pipe.set_precision(pg, pk)
        ms_alt, stages_alt = timed_steps(pipe, 1, 1)
        log(f"{pg} Gram + {pk} screen: {ms_alt:.1f} ms/step, "
            f"stages {json.dumps({k: round(v, 2) for k, v in stages_alt.items()})}")
        capture(tag)
        alts.append({"tag": tag, "prec_gram": pg, "prec_knn": pk, "ms_per_step": ms_alt,
                     "value": args.cells / (ms_alt * 1e-3), "stages_ms": {k: round(v, 3) for k, v in stages_alt.items()}})

    if args.fp32_steps > 0:
        pipe.set_precision("fp32", "fp32")
        ms32, stages32 = timed_steps(pipe, args.fp32_steps, 1)
        log(f"FP32: {ms32:.1f} ms/step, stages {json.dumps({k: round(v, 2) for k, v in stages32.items()})}")
        capture("fp32")
    pipe.set_precision(args.gram_prec, args.knn_prec)
    # T from the first upload: the CSR copied from pinned host memory into the device
    # buffers, then one step (the same pipeline), timed together
    from crest_amd._lib import check, lib
    from crest_amd.pipeline import PinnedArray
    host = [PinnedArray(a.shape, a.dtype) for a in (indptr, indices, data)]
    for h_, a in zip(host, (indptr, indices, data)):
        h_.array[...] = a
    check(lib().crest_synchronize())
    t0 = time.perf_counter()
    for h_, dbuf in zip(host, (csr.indptr, csr.indices, csr.data)):
        check(lib().crest_memcpy_htod(dbuf.ptr, h_.ptr, h_.array.nbytes))
    pipe.run()
    check(lib().crest_synchronize())
    ms_upload = (time.perf_counter() - t0) * 1e3
    upload_bytes = sum(h_.array.nbytes for h_ in host)
    del host
    log(f"upload ({upload_bytes / 1e9:.2f} GB from pinned host memory) + one step: {ms_upload:.1f} ms")
    if args.calibrate_copy:
        # a plain copy of known size: the reference reading for the HBM byte counters
        from crest_amd._lib import check, lib
        from crest_amd.pipeline import DeviceBuffer
        nb = args.calibrate_copy << 20
        a, b = DeviceBuffer(nb), DeviceBuffer(nb)
        check(lib().crest_dev_copy(b.ptr, a.ptr, nb))
        log(f"copy kernel: {nb} bytes read, {nb} bytes written")

    # T with the final device-to-host copy of the results (X_pca, kNN indices and
    # distances), and a plain device copy of 1 GiB as the measured HBM reference
    from crest_amd._lib import check, lib
    from crest_amd.pipeline import DeviceBuffer
    t0 = time.perf_counter()
    pipe.x_pca()
    pipe.knn()
    d2h_ms = (time.perf_counter() - t0) * 1e3
    nb = 1 << 30
    ca_, cb_ = DeviceBuffer(nb), DeviceBuffer(nb)
    check(lib().crest_dev_copy(cb_.ptr, ca_.ptr, nb))
    check(lib().crest_synchronize())
    t0 = time.perf_counter()
    for _ in range(5):
        check(lib().crest_dev_copy(cb_.ptr, ca_.ptr, nb))
    check(lib().crest_synchronize())
    copy_gbps = 5 * 2 * nb / (time.perf_counter() - t0) / 1e9
    del ca_, cb_

    n = args.cells
    p = pipe.kp
    screen_s = stages["knn_screen"] * 1e-3
    knn_flops = 2.0 * n * n * p
    achieved = knn_flops / screen_s
    peak = PEAK_FP32 if args.knn_prec == "fp32" else PEAK_BF16
    line = {
        "metric": METRIC, "value": n / (ms * 1e-3), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"{PREC_LABEL[args.gram_prec]} Gram + {PREC_LABEL[args.knn_prec]} kNN screen (MFMA); f64 "
                  "centring, eigensolve, projection and kNN rerank"),
        "data": "synthetic (seeded NB counts, crest_amd.synth)",
        "config": {"workload": (f"{n} cells x {args.genes} pre-selected HVGs, NB counts density "
                                f"{info['density']:.4f}, {args.clusters} clusters, seed {args.seed}; normalize_total 1e4,"
                                f" log1p, scale max_value 10, PCA {args.pcs}, exact kNN k={args.k}"),
                   "n_cells": n, "n_genes": args.genes, "nnz": info["nnz"], "n_comps": args.pcs, "k": args.k,
                   "chunk_rows": args.chunk_rows, "prec_gram": args.gram_prec, "prec_knn": args.knn_prec},
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "ms_per_step_with_d2h": ms + d2h_ms,
        "ms_from_upload": ms_upload,
        "value_from_upload": n / (ms_upload * 1e-3),
        "upload_bytes": upload_bytes,
        "hbm_copy_GBps": copy_gbps,
        "roofline": {"kernel": {"fp16": "k_knn_screen16<1,...>", "bf16x3": "k_knn_screen16<0,...>",
                                "fp32": "k_knn_screen_f32"}[args.knn_prec],
                     "bound": "mfma", "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": screen_traffic(args.knn_prec, n, p),
                     "algorithmic": "2*n^2*p flops per launch (n=%d, p=%d)" % (n, p),
                     "launch_ms": stages["knn_screen"]},
    }
    # per precision: queries whose screened kNN set the screen's error bound proved,
    # rechecked in FP32, and left unproven (near ties)
    line["knn_certificate"] = {t: gpu[f"cert_{t}"] for t in gpu["tags"]}
    if alts:
        line["alt"] = alts
    if args.fp32_steps > 0:
        line["fp32"] = {"ms_per_step": ms32, "value": n / (ms32 * 1e-3),
                        "stages_ms": {k: round(v, 3) for k, v in stages32.items()}}
        line["speedup_vs_fp32"] = ms32 / ms
        for a in alts:
            a["speedup_vs_fp32"] = ms32 / a["ms_per_step"]
    if args.streamed_cells:
        line["streamed"] = streamed_leg(args)
    if args.batch_subsets:
        line["batch"] = batched_leg(args)
    if args.c2_cells:
        line["c2"] = c2_leg(args)
    if not args.no_cpu:
        t0 = time.perf_counter()
        base, acc = cpu_baseline_leg(indptr, indices, data, args.genes, gpu, args)
        log(f"cpu baseline + accuracy leg {time.perf_counter() - t0:.1f}s")
        line["cpu_baseline"] = base
        line["accuracy"] = acc
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
