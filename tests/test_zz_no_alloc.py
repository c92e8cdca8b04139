"""crest_gram_accumulate_noalloc (scratch in a static device array, no allocation)
against crest_gram_accumulate (py.rs:270-287, kernels.rs:221-309) on the same chunks,
two calls in a row with the second smaller, in each Gram precision, and against the
oracle.  Tolerances: the two entry points sum the same products in a different
split-K order, so FP64 agrees to 1e-12 and the FP32-accumulated forms to 2e-6 of the
largest Gram entry."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import crest_amd as ca  # noqa: E402
from crest_amd import _native as N  # noqa: E402
from crest_amd import synth  # noqa: E402
from crest_amd._lib import CREST_E_VALUE, Chunk, check, lib  # noqa: E402
from crest_amd.pipeline import DeviceBuffer  # noqa: E402

N_CELLS, N_GENES = 20000, 600
TOL = {"fp64": 1e-12, "fp32": 2e-6, "fp16": 2e-6, "bf16x3": 2e-6}


@pytest.fixture(scope="module")
def data():
    ip, ix, dv, _ = synth.make_counts(N_CELLS, N_GENES, 10, density=0.08, seed=11)
    rng = np.random.default_rng(4)
    keep_c = rng.random(N_CELLS) > 0.1
    cmap = np.full(N_CELLS, -1, np.int64)
    cmap[keep_c] = np.arange(keep_c.sum())
    gmap = np.arange(N_GENES, dtype=np.int32)
    hv = np.sort(rng.choice(N_GENES, 450, replace=False))
    sub = np.full(N_GENES, -1, np.int32)
    sub[hv] = np.arange(len(hv), dtype=np.int32)
    mean = rng.uniform(0.1, 1.0, len(hv))
    sd = rng.uniform(0.5, 2.0, len(hv))
    return ip, ix, dv, cmap, gmap, sub, mean, sd


def host_ctx(ip, ix, dv, cmap, gmap, lo, hi):
    a, b = int(ip[lo]), int(ip[hi])
    return (ix[a:b], dv[a:b], cmap, gmap, 1e4, True, None, np.ascontiguousarray(ip[lo:hi + 1]), lo)


def run_noalloc(ip, ix, dv, cmap, gmap, sub, mean, sd, lo, hi, d):
    a, b = int(ip[lo]), int(ip[hi])
    bufs = [DeviceBuffer.from_array(np.ascontiguousarray(ip[lo:hi + 1] - ip[lo], np.int64)),
            DeviceBuffer.from_array(np.ascontiguousarray(ix[a:b], np.uint32)),
            DeviceBuffer.from_array(np.ascontiguousarray(dv[a:b], np.float32)),
            DeviceBuffer.from_array(cmap), DeviceBuffer.from_array(gmap), DeviceBuffer.from_array(sub),
            DeviceBuffer.from_array(mean), DeviceBuffer.from_array(sd)]
    ip_d, ix_d, dv_d, cm_d, gm_d, sub_d, mean_d, sd_d = bufs
    ch = Chunk(ix_d.ptr, dv_d.ptr, b - a, cm_d.ptr, len(cmap), gm_d.ptr, len(gmap), 1e4, 1, None, ip_d.ptr,
               hi - lo + 1, lo)
    gram = DeviceBuffer.from_array(np.zeros((d, d)))
    colsum = DeviceBuffer.from_array(np.zeros(d))
    check(lib().crest_gram_accumulate_noalloc(C.byref(ch), sub_d.ptr, len(sub), gram.ptr, colsum.ptr, d, 1,
                                              mean_d.ptr, sd_d.ptr, 10.0, None))
    check(lib().crest_stream_synchronize(None))  # the entry point leaves the wait to its caller
    return gram.to_numpy(np.float64, (d, d)), colsum.to_numpy(np.float64, (d,))


@pytest.mark.parametrize("prec", ["fp64", "fp16", "fp32", "bf16x3"])
def test_noalloc_matches_gram_accumulate(data, oracle, prec):
    ip, ix, dv, cmap, gmap, sub, mean, sd = data
    d = len(mean)
    prev = ca.get_precision()
    ca.set_precision(prec)
    try:
        for lo, hi in ((0, 12000), (12000, 15000)):  # the second call smaller
            print(f"[noalloc] {prec} cells {lo}..{hi}", flush=True)
            g_na, c_na = run_noalloc(ip, ix, dv, cmap, gmap, sub, mean, sd, lo, hi, d)
            g_op, c_op = np.zeros((d, d)), np.zeros(d)
            N.gram_accumulate(*host_ctx(ip, ix, dv, cmap, gmap, lo, hi), sub, g_op, c_op, True, mean, sd, 10.0)
            scale = np.max(np.abs(g_op))
            iu = np.triu_indices(d)
            assert np.max(np.abs(g_na[iu] - g_op[iu])) <= TOL[prec] * scale
            assert np.all(np.tril(g_na, -1) == 0.0)
            np.testing.assert_allclose(c_na, c_op, rtol=0, atol=1e-11 * np.max(np.abs(c_op)))
            if prec == "fp64":
                g_o, c_o = np.zeros((d, d)), np.zeros(d)
                oracle.gram_accumulate(*host_ctx(ip, ix, dv, cmap, gmap, lo, hi), sub, g_o, c_o, True, mean, sd,
                                       10.0, threads=8)
                assert np.max(np.abs(g_na[iu] - g_o[iu])) <= 1e-12 * scale
    finally:
        ca.set_precision(prev)


def test_noalloc_rejects_beyond_scratch():
    """d above the scratch's 2,048 HVGs is a ValueError raised before any GPU work."""
    ip_d = DeviceBuffer.from_array(np.zeros(2, np.int64))
    out = DeviceBuffer(256)
    ch = Chunk(None, None, 0, None, 0, None, 0, 1e4, 1, None, ip_d.ptr, 2, 0)
    rc = lib().crest_gram_accumulate_noalloc(C.byref(ch), None, 0, out.ptr, out.ptr, 2049, 0, None, None, 0.0, None)
    assert rc == CREST_E_VALUE
    with pytest.raises(ValueError, match="no-allocation scratch"):
        check(rc)
