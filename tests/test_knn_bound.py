"""CPU checks of the kNN screen certificate's error bound (crest_knn_screen_bound).

The GPU screens compute |c|^2 + |q|^2 - 2 c.q from rounded operands (FP16 of the
prescaled centred rows, split-BF16 hi/lo, or FP32) with FP32 accumulation.  The
certificate relies on |screen - s^2 |c - q|^2| <= c1 a b + c2 (a + b)^2 + c3 (a + b)
+ c4 (a = s|q'|, b = s|c'|).  Here the operands are formed with the same roundings
(numpy float16 and float32 round to nearest even, split-BF16 by the same bit
arithmetic), the screen value is accumulated in FP32, and the library's own
coefficients must cover every pair of adversarial sets: rows far from the mean,
near-duplicates, columns whose scaled values are FP16 subnormals.  A bound with the
leading coefficient weakened must fail on the same pairs.
"""
import ctypes as C

import numpy as np
import pytest

from crest_amd import _lib

PREC = {"gemm": -1, "fp32": 1, "bf16x3": 2, "fp16": 3}


def _bound(prec, p):
    coef = np.zeros(4, np.float64)
    kpad = C.c_int32(0)
    _lib.check(_lib.lib().crest_knn_screen_bound(PREC[prec], p, coef.ctypes.data, C.byref(kpad)))
    return coef, kpad.value


def _sumsq_f32(v):
    nn = np.zeros(v.shape[0], np.float32)
    for f in range(v.shape[1]):
        sq = (v[:, f] * v[:, f]).astype(np.float32)
        nn = (nn + sq).astype(np.float32)
    return nn


def _acc_f32(terms):
    s = np.zeros(terms[0].shape, np.float32)
    for t in terms:
        s = (s + t.astype(np.float32)).astype(np.float32)
    return s


def _bf16(v):
    u = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _prescale(x, mean):
    """Power of two s with s^2 max|x'|^2 <= 2^14 (the FP16 operands' prescale)."""
    mx = float(_sumsq_f32((x - mean).astype(np.float32)).max())
    _, ex = np.frexp(mx)
    e = 14 - int(ex)
    e = e // 2
    return float(2.0 ** max(-60, min(60, e)))


def _screen(prec, X, mean, ic, iq):
    """Screen values of pairs (ic[j], iq[j]) of rows of X, and the prescale."""
    s = _prescale(X, mean) if prec == "fp16" else 1.0
    V = ((X - mean).astype(np.float32) * np.float32(s)).astype(np.float32)
    nn = _sumsq_f32(V)
    Vc, Vq = V[ic], V[iq]
    p = X.shape[1]
    if prec == "fp16":
        H = V.astype(np.float16).astype(np.float32)
        nh = nn.astype(np.float16).astype(np.float32)
        nl = (nn - nh).astype(np.float32).astype(np.float16).astype(np.float32)
        terms = [H[ic][:, f] * (-2.0 * H[iq][:, f]) for f in range(p)]
        terms += [nh[ic], nl[ic], nh[iq], nl[iq]]
    elif prec == "bf16x3":
        H = _bf16(V)
        Lo = _bf16((V - H).astype(np.float32))
        nh = _bf16(nn)
        nl = _bf16((nn - nh).astype(np.float32))
        terms = []
        for f in range(p):
            terms.append(H[ic][:, f] * (-2.0 * H[iq][:, f]))
            terms.append(H[ic][:, f] * (-2.0 * Lo[iq][:, f]))
            terms.append(Lo[ic][:, f] * (-2.0 * H[iq][:, f]))
        terms += [nh[ic], nl[ic], nh[iq], nl[iq]]
    elif prec == "gemm":
        # rows centred in FP32: fl(fl(|q'|^2 + |c'|^2) - 2 fl(q'.c')), clamped at 0
        X32 = V
        n2 = _sumsq_f32(X32)
        dot = _acc_f32([(X32[ic][:, f] * X32[iq][:, f]).astype(np.float32) for f in range(p)])
        key = ((n2[ic] + n2[iq]).astype(np.float32) - (np.float32(2.0) * dot)).astype(np.float32)
        return np.maximum(key, 0).astype(np.float64), 1.0
    else:
        terms = [(Vc[:, f] * (-2.0 * Vq[:, f])).astype(np.float32) for f in range(p)]
        terms += [nn[ic], nn[iq]]
    return _acc_f32(terms).astype(np.float64), s


def _ratio(prec, X, ic, iq, coef):
    mean = X.astype(np.float64).mean(0).astype(np.float32)
    scr, s = _screen(prec, X, mean, ic, iq)
    X64, m64 = X.astype(np.float64), mean.astype(np.float64)
    D = ((X64[ic] - X64[iq]) ** 2).sum(1)
    a = s * np.sqrt(((X64[iq] - m64) ** 2).sum(1)) * (1 + 1e-6)
    b = s * np.sqrt(((X64[ic] - m64) ** 2).sum(1)) * (1 + 1e-6)
    E = coef[0] * a * b + coef[1] * (a + b) ** 2 + coef[2] * (a + b) + coef[3]
    return np.abs(scr - s * s * D) / E


def _sets(p, seed):
    rng = np.random.default_rng(seed)
    n = 4000
    out = []
    # far from the mean, neighbours close: two tight clusters at +-R
    for R in (30.0, 300.0, 3e4):
        X = rng.normal(size=(n, p)) + np.where(np.arange(n)[:, None] % 2 == 0, R, -R) * (np.arange(p) == 0)
        out.append(X.astype(np.float32))
    # near-duplicates of a few base rows
    base = rng.normal(size=(40, p)) * 5
    out.append((base[np.arange(n) % 40] + rng.normal(size=(n, p)) * 1e-4).astype(np.float32))
    # one dominant column: the other columns' scaled values are FP16 subnormals
    X = rng.normal(size=(n, p)) * 1e-3
    X[:, 0] = rng.normal(size=n) * 2e3
    out.append(X.astype(np.float32))
    # tiny overall scale (prescale up) and huge (prescale down)
    out.append((rng.normal(size=(n, p)) * 1e-12).astype(np.float32))
    out.append((rng.normal(size=(n, p)) * 1e12).astype(np.float32))
    return out


def _pairs(X, rng):
    n = X.shape[0]
    iq = rng.integers(0, n, 20000)
    # nearby rows (same cluster / base row / parity) and random rows
    ic = np.where(np.arange(iq.size) % 2 == 0, (iq + 2 * rng.integers(1, 50, iq.size)) % n, rng.integers(0, n, iq.size))
    keep = ic != iq
    return ic[keep], iq[keep]


@pytest.mark.parametrize("prec", ["fp16", "bf16x3", "fp32", "gemm"])
@pytest.mark.parametrize("p", [7, 50, 200])
def test_screen_error_within_bound(prec, p):
    coef, kpad = _bound(prec, p)
    if kpad < 0:
        assert not coef.any()
        pytest.skip(f"no {prec} register screen for {p} columns")
    assert coef[1] > 0
    rng = np.random.default_rng(p)
    worst = 0.0
    for X in _sets(p, p):
        ic, iq = _pairs(X, rng)
        worst = max(worst, float(_ratio(prec, X, ic, iq, coef).max()))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("prec", ["fp16", "bf16x3"])
def test_weakened_bound_fails(prec):
    """The leading (operand rounding) coefficient is needed: a quarter of it is exceeded."""
    p = 50
    coef, _ = _bound(prec, p)
    weak = coef.copy()
    weak[0] *= 0.25
    weak[1] = weak[2] = weak[3] = 0.0
    rng = np.random.default_rng(1)
    X = _sets(p, 3)[1]
    ic, iq = _pairs(X, rng)
    assert float(_ratio(prec, X, ic, iq, weak).max()) > 1.0


def test_bound_coefficients_shape():
    c16, k16 = _bound("fp16", 50)
    cbf, kbf = _bound("bf16x3", 50)
    c32, k32 = _bound("fp32", 50)
    assert k16 == 64 and kbf == 160 and k32 == 56
    assert c32[0] == 0.0 and cbf[0] < c16[0]
    assert _bound("fp32", 500)[1] == -1
    cg, kg = _bound("gemm", 500)
    assert kg == 500 and cg[0] == 0.0 and cg[1] > c32[1]
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().crest_knn_screen_bound(9, 5, np.zeros(4).ctypes.data, C.byref(C.c_int32())))
