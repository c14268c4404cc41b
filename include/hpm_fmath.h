/*
 * hpm_fmath.h — elementary functions built only from IEEE-754 basic
 * operations (+ - * / sqrt fma, rint), so the host and gfx950 return the same
 * bits for the same input. The reference compiles with --use_fast_math
 * (HPM-MVS/CMakeLists.txt:14); its exp/sin/cos/acos are hardware
 * approximations that cannot be reproduced, so these functions define the
 * arithmetic instead. They are checked against the host libm to a few ulp.
 *
 *  hpm_expf : Cody & Waite (1980) range reduction x = n*ln2 + r with a
 *             two-part ln2, degree-7 Taylor polynomial on |r| <= ln2/2,
 *             scaling by 2^n in two exact steps (keeps subnormal results
 *             correctly rounded once).
 *  hpm_sinf / hpm_cosf : Cody-Waite reduction by pi/2 with a three-part
 *             constant (the Cephes split), Taylor polynomials on [-pi/4, pi/4].
 *  hpm_acosf: the Cephes asinf scheme (S. L. Moshier, "Methods and Programs
 *             for Mathematical Functions", 1989): a degree-4 polynomial in
 *             z = x^2 for |x| <= 0.5, and asin(x) = pi/2 - 2 asin(sqrt((1-x)/2))
 *             above. |x| > 1 gives NaN, as acosf does.
 *
 * Every file that uses these must be compiled without FP contraction
 * (-ffp-contract=off) and without fast-math, and with correctly rounded
 * FP32 division and square root on the GPU.
 */
#ifndef HPM_FMATH_H_
#define HPM_FMATH_H_

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define HPM_HD_F __host__ __device__
#else
#define HPM_HD_F
#endif

static inline HPM_HD_F float hpm_bits_to_float(uint32_t u)
{
    float f;
    __builtin_memcpy(&f, &u, sizeof(f));
    return f;
}

static inline HPM_HD_F float hpm_qnan(void)
{
    return hpm_bits_to_float(0x7fc00000u);
}

/* Branch-free: the core runs on x clamped to [-103.97, 88.72] (x itself
 * inside), and NaN, overflow and underflow are selected afterwards, so the
 * GPU schedules calls side by side instead of behind per-call branches. */
static inline HPM_HD_F float hpm_expf(float x)
{
    const float lo = -103.972084045410156f, hi = 88.72283935546875f;
    const float xc = fminf(fmaxf(x, lo), hi);
    const float n = rintf(xc * 1.44269502162933349609f);
    float r = fmaf(n, -0.693145751953125f, xc);
    r = fmaf(n, -1.428606765330187045e-06f, r);
    float p = 1.98412698412698413e-04f;
    p = fmaf(p, r, 1.38888888888888889e-03f);
    p = fmaf(p, r, 8.33333333333333333e-03f);
    p = fmaf(p, r, 4.16666666666666667e-02f);
    p = fmaf(p, r, 1.66666666666666667e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const int ni = (int)n;
    const int n1 = ni / 2;
    const int n2 = ni - n1;
    const float s1 = hpm_bits_to_float((uint32_t)(n1 + 127) << 23);
    const float s2 = hpm_bits_to_float((uint32_t)(n2 + 127) << 23);
    float y = (p * s1) * s2;
    y = x > hi ? hpm_bits_to_float(0x7f800000u) : y;
    y = x < lo ? 0.0f : y;
    return x == x ? y : x;
}

static inline HPM_HD_F float hpm_sin_poly(float r)
{
    const float z = r * r;
    float p = fmaf(z, 2.75573192239858907e-06f, -1.98412698412698413e-04f);
    p = fmaf(z, p, 8.33333333333333333e-03f);
    p = fmaf(z, p, -1.66666666666666667e-01f);
    return fmaf(r * z, p, r);
}

static inline HPM_HD_F float hpm_cos_poly(float r)
{
    const float z = r * r;
    float p = fmaf(z, -2.75573192239858907e-07f, 2.48015873015873016e-05f);
    p = fmaf(z, p, -1.38888888888888889e-03f);
    p = fmaf(z, p, 4.16666666666666667e-02f);
    p = fmaf(z, p, -0.5f);
    return fmaf(z, p, 1.0f);
}

/* quadrant and reduced argument of x / (pi/2) */
static inline HPM_HD_F float hpm_reduce_pio2(float x, int* quadrant)
{
    const float n = rintf(x * 0.636619772367581343f);
    float r = fmaf(n, -1.5703125f, x);
    r = fmaf(n, -4.837512969970703125e-4f, r);
    r = fmaf(n, -7.54978995489188216e-8f, r);
    *quadrant = ((int)n) & 3;
    return r;
}

static inline HPM_HD_F float hpm_sinf(float x)
{
    if (!(fabsf(x) <= 1.0e5f)) {
        return hpm_qnan();
    }
    int q;
    const float r = hpm_reduce_pio2(x, &q);
    const float s = hpm_sin_poly(r);
    const float c = hpm_cos_poly(r);
    return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}

static inline HPM_HD_F float hpm_cosf(float x)
{
    if (!(fabsf(x) <= 1.0e5f)) {
        return hpm_qnan();
    }
    int q;
    const float r = hpm_reduce_pio2(x, &q);
    const float s = hpm_sin_poly(r);
    const float c = hpm_cos_poly(r);
    return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

static inline HPM_HD_F float hpm_asin_poly(float z)
{
    float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = fmaf(p, z, 4.5470025998e-2f);
    p = fmaf(p, z, 7.4953002686e-2f);
    p = fmaf(p, z, 1.6666752422e-1f);
    return p;
}

static inline HPM_HD_F float hpm_acosf(float x)
{
    const float a = fabsf(x);
    if (!(a <= 1.0f)) {
        return hpm_qnan();
    }
    const float pio2_hi = 1.57079637050628662109375f;
    const float pio2_lo = -4.37113900018624283e-8f;
    const float pi_hi = 3.14159274101257324219f;
    const float pi_lo = -8.74227800037248566e-8f;
    if (a <= 0.5f) {
        const float z = x * x;
        const float asin_x = fmaf(x * z, hpm_asin_poly(z), x);
        return (pio2_hi - asin_x) + pio2_lo;
    }
    const float z = 0.5f * (1.0f - a);
    const float s = sqrtf(z);
    const float asin_s = fmaf(s * z, hpm_asin_poly(z), s);
    const float r = 2.0f * asin_s;
    if (x > 0.0f) {
        return r;
    }
    return (pi_hi - r) + pi_lo;
}

#endif /* HPM_FMATH_H_ */
