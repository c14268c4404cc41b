/*
 * hpm_half.h — FP32 <-> FP16 conversion of source-image texels.
 *
 * The padded source pool on the device holds FP16 texels: half the bytes of
 * every NCC gather, and a footprint that fits the vector L1. Grey values of
 * 8-bit images (the only images the reference reads, HPM.cpp:492-511) are
 * exact in FP16; other values (images resized to max_image_size,
 * HPM.cpp:524-548) are sampled as their nearest FP16 value, by the device and
 * by the CPU restatement alike. Round to nearest, ties to even, as IEEE 754
 * binary16 conversion; values beyond the binary16 range become infinities.
 */
#ifndef HPM_HALF_H
#define HPM_HALF_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint16_t hpm_f16_bits(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) { /* infinity, NaN (quiet) */
        return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    }
    if (ax >= 0x477ff000u) { /* >= 65520: rounds to infinity */
        return (uint16_t)(sign | 0x7c00u);
    }
    if (ax < 0x38800000u) { /* below 2^-14: binary16 subnormal or zero */
        if (ax <= 0x33000000u) { /* <= 2^-25: ties to the even zero */
            return (uint16_t)sign;
        }
        const uint32_t e = ax >> 23;
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t s = 126u - e; /* value / 2^-24 = m >> s */
        uint32_t k = m >> s;
        const uint32_t rem = m & ((1u << s) - 1u), half = 1u << (s - 1u);
        if (rem > half || (rem == half && (k & 1u))) {
            ++k;
        }
        return (uint16_t)(sign | k);
    }
    uint32_t h = (ax >> 13) - ((127u - 15u) << 10);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) {
        ++h; /* a carry into the exponent is the correct rounding */
    }
    return (uint16_t)(sign | h);
}

static inline float hpm_f16_value(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0u) {
        const float v = (float)m * 5.9604644775390625e-8f; /* m * 2^-24, exact */
        memcpy(&x, &v, 4);
        x |= sign;
    } else if (e == 31u) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* the value a source texel is sampled as */
static inline float hpm_round_f16(float f)
{
    return hpm_f16_value(hpm_f16_bits(f));
}

/* The bilinear coefficients of a 2x2 texel footprint, as FP16: with the
 * texels t00 = (x, y), t10 = (x + 1, y), t01 = (x, y + 1), t11 = (x + 1,
 * y + 1) at their sampled values (hpm_round_f16: the arguments must be
 * FP16 values already),
 *   c[0] = t00, c[1] = t10 - t00, c[2] = t01 - t00,
 *   c[3] = (t11 - t01) - (t10 - t00),
 * differences in FP32 then rounded to FP16. For 8-bit images every
 * coefficient is an integer of magnitude <= 510, exact in FP16. A sample at
 * fractions (a, b) is then hpm_bilinear_eval: three fused multiply-adds. */
static inline void hpm_bilinear_coeffs(float t00, float t10, float t01, float t11, uint16_t c[4])
{
    c[0] = hpm_f16_bits(t00);
    c[1] = hpm_f16_bits(t10 - t00);
    c[2] = hpm_f16_bits(t01 - t00);
    c[3] = hpm_f16_bits((t11 - t01) - (t10 - t00));
}

/* the sample at fractions (a, b) of a footprint's coefficients:
 * (c0 + a c1) + b (c2 + a c3), each bracket and the sum one fused
 * multiply-add */
static inline float hpm_bilinear_eval(const uint16_t c[4], float a, float b)
{
    const float v0 = fmaf(a, hpm_f16_value(c[1]), hpm_f16_value(c[0]));
    const float v = fmaf(a, hpm_f16_value(c[3]), hpm_f16_value(c[2]));
    return fmaf(b, v, v0);
}

#endif
