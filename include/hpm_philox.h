/*
 * hpm_philox.h — counter-based Philox4x32-10 random numbers, identical on the
 * host (gcc / clang) and on gfx950.
 *
 * Written from J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11: a 4x32-bit counter
 * and a 2x32-bit key go through 10 rounds of two 32x32->64 multiplies and a
 * Weyl key schedule.
 *
 * The reference draws from cuRAND XORWOW states seeded with clock64()
 * (HPM-MVS/HPM.cu:521), which is not reproducible. Here every draw is a pure
 * function of (seed, pixel, stream, salt, draw index), so the GPU and the CPU
 * oracle consume exactly the same numbers in the same order.
 *
 * Counter layout of one draw sequence:
 *   ctr = { draw_index / 4, pixel_linear_index, stream_id, salt }, key = seed,
 *   word = output[draw_index % 4].
 * stream_id separates launches (RunPatchMatch call, init / iteration / colour),
 * salt separates reference views.
 */
#ifndef HPM_PHILOX_H_
#define HPM_PHILOX_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define HPM_HD __host__ __device__
#else
#define HPM_HD
#endif

#define HPM_PHILOX_M0 0xD2511F53u
#define HPM_PHILOX_M1 0xCD9E8D57u
#define HPM_PHILOX_W0 0x9E3779B9u
#define HPM_PHILOX_W1 0xBB67AE85u

static inline HPM_HD void hpm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)HPM_PHILOX_M0 * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)HPM_PHILOX_M1 * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += HPM_PHILOX_W0;
        k1 += HPM_PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Uniform float in (0, 1]: the top 24 bits plus one, times 2^-24. Every
 * value is exact in FP32; 0 is never returned and 1 is (as with the
 * reference's curand_uniform, whose range is also (0, 1]). */
static inline HPM_HD float hpm_u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; /* 2^-24 */
}

/* One pixel's draw sequence within one launch. */
typedef struct hpm_rng {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    uint32_t left; /* unused words left in buf */
} hpm_rng;

static inline HPM_HD void hpm_rng_init(hpm_rng* r, uint32_t seed_lo, uint32_t seed_hi,
                                       uint32_t pixel, uint32_t stream, uint32_t salt)
{
    r->key[0] = seed_lo;
    r->key[1] = seed_hi;
    r->ctr[0] = 0u;
    r->ctr[1] = pixel;
    r->ctr[2] = stream;
    r->ctr[3] = salt;
    r->left = 0u;
}

static inline HPM_HD uint32_t hpm_rng_next_u32(hpm_rng* r)
{
    if (r->left == 0u) {
        hpm_philox4x32_10(r->ctr, r->key, r->buf);
        r->ctr[0] += 1u;
        r->left = 4u;
    }
    const uint32_t i = 4u - r->left;
    r->left -= 1u;
    /* select instead of a dynamic index so the buffer stays in registers */
    return i == 0u ? r->buf[0] : (i == 1u ? r->buf[1] : (i == 2u ? r->buf[2] : r->buf[3]));
}

static inline HPM_HD float hpm_rng_uniform(hpm_rng* r)
{
    return hpm_u01(hpm_rng_next_u32(r));
}

/* Stream ids: one per kernel launch of one RunPatchMatch call. */
static inline HPM_HD uint32_t hpm_stream_init(uint32_t call_index)
{
    return call_index * 16u;
}
static inline HPM_HD uint32_t hpm_stream_update(uint32_t call_index, int iter, int colour)
{
    return call_index * 16u + 1u + (uint32_t)(2 * iter + colour);
}

#endif /* HPM_PHILOX_H_ */
