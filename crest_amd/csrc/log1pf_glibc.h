/*
 * log1pf with the rounding of glibc 2.35's log1pf, for the lazy transform.
 *
 * The reference computes the log1p of the normalised value in f32 with Rust's
 * f32::ln_1p (src/kernels.rs:81), which calls the platform C library's log1pf.  On
 * glibc 2.35 (x86-64) that function is not correctly rounded: over all floats >= -1
 * it differs from the correctly rounded result on 9,149,209 inputs.  To give results
 * bit-identical to the reference, the device evaluates the same algorithm in IEEE
 * single precision, step for step: the fdlibm log1p scheme (Sun Microsystems, 1993,
 * as published in fdlibm / FreeBSD msun and used by glibc's flt-32 port):
 *
 *   1 + x = 2^k (1 + f),  sqrt(2)/2 <= 1 + f < sqrt(2);
 *   c     = the rounding error of the float sum 1 + x, divided by 1 + x;
 *   s     = f / (2 + f),  log(1 + f) = f - (f^2/2 - s (f^2/2 + R(s^2)))
 *           with R(z) = z (L1 + z (L2 + ... + z L7)), a minimax fit to the
 *           odd series 2 s^3/3 + 2 s^5/5 + ...;
 *   log1p(x) = k ln2_hi - ((f^2/2 - (s (f^2/2 + R) + (k ln2_lo + c))) - f).
 *
 * Branch thresholds are glibc's (they differ from BSD's float port): the short series
 * x - x^2/2 below 2^-29, the identity below 2^-54, the no-reduction range
 * -0.29289 < x < 0.41421 (bit patterns 0xbe95f61f / 0x3ed413d7), and the
 * renormalisation of the mantissa at 0x3504f7.  Checked bit for bit against glibc's
 * log1pf on every float >= -1 (tests/test_log1pf.py).
 *
 * The includer defines CREST_LOG1PF_FN (the function qualifiers) before including.
 * Evaluation must be IEEE single precision with no contraction of multiply-adds and
 * correctly rounded division (hipcc's defaults with -ffp-contract=off).
 */
#ifndef CREST_LOG1PF_GLIBC_H
#define CREST_LOG1PF_GLIBC_H

#include <stdint.h>

CREST_LOG1PF_FN float crest_f32_from_bits(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
CREST_LOG1PF_FN uint32_t crest_bits_of_f32(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

CREST_LOG1PF_FN float crest_log1pf_glibc(float x) {
    /* ln 2 split so that k * ln2_hi is exact for |k| < 2^11, and the fit's coefficients */
    const float ln2_hi = crest_f32_from_bits(0x3f317180u), ln2_lo = crest_f32_from_bits(0x3717f7d1u);
    const float L1 = crest_f32_from_bits(0x3f2aaaabu), L2 = crest_f32_from_bits(0x3ecccccdu),
                L3 = crest_f32_from_bits(0x3e924925u), L4 = crest_f32_from_bits(0x3e638e29u),
                L5 = crest_f32_from_bits(0x3e3a3325u), L6 = crest_f32_from_bits(0x3e1cd04fu),
                L7 = crest_f32_from_bits(0x3e178897u);
    const int32_t hx = (int32_t)crest_bits_of_f32(x);
    const int32_t ax = hx & 0x7fffffff;
    int32_t k = 1;      /* power of two taken out of 1 + x (0: no reduction)  */
    int32_t hu = 0;     /* mantissa bits of the reduced 1 + f (0: |f| tiny)     */
    float f = 0.0f, c = 0.0f;

    if (hx < 0x3ed413d7) { /* x < 0.41421 (or negative) */
        if (ax >= 0x3f800000) { /* x <= -1 */
            if (x == -1.0f) return -__builtin_inff();
            return __builtin_nanf("");
        }
        if (ax < 0x31000000) { /* |x| < 2^-29: x - x^2/2, or x itself below 2^-54 */
            if (ax < 0x24800000) return x;
            return x - x * x * 0.5f;
        }
        if (hx > 0 || hx <= (int32_t)0xbe95f61f) { /* -0.29289 <= x < 0.41421: f = x */
            k = 0;
            f = x;
            hu = 1;
        }
    }
    if (hx >= 0x7f800000) return x + x; /* +inf, NaN */
    if (k != 0) {
        float u;
        if (hx < 0x5a000000) { /* x < 2^53: 1 + x rounds; keep its error in c */
            u = 1.0f + x;
            hu = (int32_t)crest_bits_of_f32(u);
            k = (hu >> 23) - 127;
            c = (k > 0) ? 1.0f - (u - x) : x - (u - 1.0f);
            c = c / u;
        } else {
            u = x;
            hu = (int32_t)crest_bits_of_f32(u);
            k = (hu >> 23) - 127;
            c = 0.0f;
        }
        hu &= 0x007fffff;
        if (hu < 0x3504f7) { /* mantissa below sqrt(2): 1 + f in [1, sqrt(2)) */
            u = crest_f32_from_bits((uint32_t)hu | 0x3f800000u);
        } else { /* else halve it into [sqrt(2)/2, 1) */
            k += 1;
            u = crest_f32_from_bits((uint32_t)hu | 0x3f000000u);
            hu = (0x00800000 - hu) >> 2;
        }
        f = u - 1.0f;
    }
    const float hfsq = 0.5f * f * f;
    if (hu == 0) { /* |f| < 2^-20 */
        if (f == 0.0f) {
            if (k == 0) return 0.0f;
            c += (float)k * ln2_lo;
            return (float)k * ln2_hi + c;
        }
        const float R = hfsq * (1.0f - 0.66666666666666666f * f);
        if (k == 0) return f - R;
        return (float)k * ln2_hi - ((R - ((float)k * ln2_lo + c)) - f);
    }
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float R = z * (L1 + z * (L2 + z * (L3 + z * (L4 + z * (L5 + z * (L6 + z * L7))))));
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return (float)k * ln2_hi - ((hfsq - (s * (hfsq + R) + ((float)k * ln2_lo + c))) - f);
}

#endif /* CREST_LOG1PF_GLIBC_H */
