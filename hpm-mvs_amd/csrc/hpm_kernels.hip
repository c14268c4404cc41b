/*
 * hpm_kernels.hip — CDNA4 (gfx950) kernels of the per-reference-view
 * PatchMatch: random/prior/GC initialisation, red/black checkerboard
 * propagation with NESP sampling, joint view selection and refinement,
 * plane->depth finalisation, the red/black median filter, fixed-hypothesis
 * cost vectors and joint bilateral upsampling.
 *
 * Reference semantics: HPM-MVS/HPM.cu (cited per function). The arithmetic
 * is the shared FP32 definition documented in oracle/hpm_oracle.cpp's
 * header: the same expression order, explicit fmaf where used, correctly
 * rounded division and square root, hpm_fmath.h elementary functions and
 * Philox random numbers, so results can be compared bit for bit with the
 * CPU restatement. Build with -ffp-contract=off.
 *
 * Mapping: one lane per pixel of the active colour, 64-wide waves along x.
 * The 36 reference taps, their bilateral weights and reference sums stay in
 * VGPRs for the whole pixel update; source taps are bilinear gathers from
 * the FP32 image pool (cache resident: a 1600x1066 view set is ~75 MB).
 */
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "hpm_device.h"
#include "hpm_fmath.h"
#include "hpm_philox.h"

#define HPM_NTAP 36 /* 11x11 window at stride 2 (HPM.cu:347-356) */

/* Waves per SIMD the NCC kernels are register-budgeted for: 2 leaves 256
 * VGPRs, room for the 72 reference weights plus ~25 gathers in flight. */
#ifndef HPM_NCC_WAVES
#define HPM_NCC_WAVES 2
#endif


namespace {

constexpr double kPi = 3.14159265358979323846;

/* The view table is written by the host before a launch and only read by
 * the kernels: reading it through the constant address space lets the
 * compiler use scalar loads wherever the view index is wave-uniform. */
typedef const __attribute__((address_space(4))) hpm_dev_view* hpm_cview;

__device__ __forceinline__ hpm_cview cview(const hpm_dev_problem& pb, int v)
{
    return (hpm_cview)pb.views + v;
}

__device__ __forceinline__ hpm_camera camera(const hpm_dev_problem& pb, int v)
{
    const hpm_cview sv = cview(pb, v);
    hpm_camera c;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        c.K[i] = sv->cam.K[i];
        c.R[i] = sv->cam.R[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        c.t[i] = sv->cam.t[i];
    }
    c.height = sv->cam.height;
    c.width = sv->cam.width;
    c.depth_min = sv->cam.depth_min;
    c.depth_max = sv->cam.depth_max;
    return c;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ float4 ld4(const float* a, int i)
{
    return reinterpret_cast<const float4*>(a)[i];
}

__device__ __forceinline__ void st4(float* a, int i, float4 v)
{
    reinterpret_cast<float4*>(a)[i] = v;
}

__device__ __forceinline__ float texel(const float* img, int W, int H, int x, int y)
{
    return img[(size_t)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1)];
}

/* Correctly rounded square root, quotient and reciprocal for operands that
 * need no range scaling: the hardware's approximations refined exactly as
 * the general (scaling, special-case fixing) instruction sequences refine
 * them, without the steps that only act on denormal, infinite, zero or
 * extreme-exponent operands. For operands in the stated ranges the results
 * are the IEEE ones (checked against the host's on the GPU). */

/* sqrt(x) for normal positive finite x: the hardware root s, then whichever
 * of s - 1 ulp, s, s + 1 ulp the exact residuals x - t s select */
__device__ __forceinline__ float sqrt_rn(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float t = fmaf(-sd, s, x) <= 0.0f ? sd : s;
    return fmaf(-su, s, x) > 0.0f ? su : t;
}

/* n / d for normal d and |n| / |d| far from the exponent range's ends (n may
 * be zero): reciprocal, one Newton step, the quotient and two exact
 * residual corrections */
__device__ __forceinline__ float div_rn(float n, float d)
{
    float r = __builtin_amdgcn_rcpf(d);
    r = fmaf(fmaf(-d, r, 1.0f), r, r);
    float q = n * r;
    q = fmaf(fmaf(-d, q, n), r, q);
    return fmaf(fmaf(-d, q, n), r, q);
}

/* 1 / d for normal d of moderate exponent: reciprocal, two Newton steps and
 * an exact residual correction. A zero d gives NaN (the general sequence
 * gives an infinity); both make every warp built from it fail its bounds
 * test, so a hypothesis scores the same. */
__device__ __forceinline__ double rcp_rn(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(fma(-d, r, 1.0), r, r);
}

/* q = Kr^-T n / w of one hypothesis (double); the homography of source
 * view v is then H = A_v - b_v q^T (HPM.cu:227-287) */
__device__ __forceinline__ void plane_q(const hpm_dev_problem& pb, float4 n, double q[3])
{
    const double iw = rcp_rn((double)n.w);
    const double nx = n.x, ny = n.y, nz = n.z;
    q[0] = (nx * pb.rk_ifx) * iw;
    q[1] = (ny * pb.rk_ify) * iw;
    q[2] = ((nz - (nx * pb.rk_cx) * pb.rk_ifx) - (ny * pb.rk_cy) * pb.rk_ify) * iw;
}

/* Warp of one (hypothesis, source view) around reference pixel p: the
 * warped centre (ComputeCorrespondingPoint, HPM.cu:289-296) in double, split
 * into an integer texel and a float fraction, plus the float coefficients of
 * the tap offsets s(p + d) - s0 = (g d) / (1 + e . d). */
struct Warp {
    uint32_t base8;       /* byte offset of entry (ix0, iy0) in the padded pool (< 2^29 entries) */
    float fx0, fy0;
    float ex, ey, gxx, gxy, gyx, gyy;
    float lox, hix, loy, hiy; /* clamp of the offsets: source coordinate in [-1, W] x [-1, H] */
    bool small;           /* |e . d| <= 0.01 over the window: cubic reciprocal */
    bool tiny;            /* |e . d| <= 0.002 over the window: linear reciprocal */
    bool inner;           /* tiny, and every tap lies inside the clamp: the clamp is the identity */
};

/* false when the warped centre leaves the source image (HPM.cu:333-335) */
__device__ __forceinline__ bool make_warp(hpm_cview sv, const double q[3], int px, int py, Warp* w)
{
    const int SW = sv->cam.width, SH = sv->cam.height;
    double H[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            H[3 * r + c] = fma(-sv->b[r], q[c], sv->A[3 * r + c]);
        }
    }
    const double X = (double)px, Y = (double)py;
    const double hx = fma(H[0], X, fma(H[1], Y, H[2]));
    const double hy = fma(H[3], X, fma(H[4], Y, H[5]));
    const double hz = fma(H[6], X, fma(H[7], Y, H[8]));
    const double icz = rcp_rn(hz);
    const double sx0 = hx * icz;
    const double sy0 = hy * icz;
    if (!(sx0 >= 0.0 && sx0 < (double)SW && sy0 >= 0.0 && sy0 < (double)SH)) {
        return false;
    }
    const double fx = floor(sx0), fy = floor(sy0);
    const int ix0 = (int)fx, iy0 = (int)fy;
    w->base8 = ((uint32_t)sv->pad_off + (uint32_t)((iy0 + 1) * sv->pitch) + (uint32_t)(ix0 + 1)) * 8u;
    w->fx0 = (float)(sx0 - fx);
    w->fy0 = (float)(sy0 - fy);
    w->ex = (float)(H[6] * icz);
    w->ey = (float)(H[7] * icz);
    w->gxx = (float)(fma(-sx0, H[6], H[0]) * icz);
    w->gxy = (float)(fma(-sx0, H[7], H[1]) * icz);
    w->gyx = (float)(fma(-sy0, H[6], H[3]) * icz);
    w->gyy = (float)(fma(-sy0, H[7], H[4]) * icz);
    w->lox = -1.0f - (float)ix0;
    w->hix = (float)(SW - ix0);
    w->loy = -1.0f - (float)iy0;
    w->hiy = (float)(SH - iy0);
    w->small = (fabsf(w->ex) + fabsf(w->ey)) * 5.0f <= 0.01f;
    w->tiny = (fabsf(w->ex) + fabsf(w->ey)) * 5.0f <= 0.002f;
    /* |tap offset| <= 5 (|g.x| + |g.y|) / (1 - 0.01) for small warps; 1.0102
     * covers the maxima of the cubic (1.010101) and the linear form (1.002)
     * and the FP32 roundings, 1e-3 px the roundings of these bounds */
    const float bx = fmaf(5.0f * (fabsf(w->gxx) + fabsf(w->gxy)), 1.0102f, 1e-3f);
    const float by = fmaf(5.0f * (fabsf(w->gyx) + fabsf(w->gyy)), 1.0102f, 1e-3f);
    w->inner = w->tiny && w->fx0 - bx >= w->lox && w->fx0 + bx <= w->hix && w->fy0 - by >= w->loy &&
               w->fy0 + by <= w->hiy;
    return true;
}

/* Paths of the window sums, chosen per wave: every active lane has an
 * inner warp (the linear reciprocal, no clamp), a small warp (per lane the
 * linear or the cubic, with the clamp), or any warp (per lane the linear,
 * the cubic or the quotient, with the clamp). The three give the same
 * values per lane. */
enum : int { kPathLinear = 0, kPathSmall = 1, kPathAny = 2 };

/* The dx products of one window column (x offset dx), shared by its taps. */
struct Column {
    float ex, gx, gy; /* ex dx, gxx dx, gyx dx */
    float omex;       /* 1 - ex dx, one rounding: the linear reciprocal's column term */
};

template <int kPath>
__device__ __forceinline__ Column column(const Warp& w, float dx)
{
    Column c;
    if (kPath != kPathLinear) {
        c.ex = w.ex * dx;
    }
    c.omex = fmaf(-w.ex, dx, 1.0f);
    c.gx = w.gxx * dx;
    c.gy = w.gyx * dx;
    return c;
}

/* floor(x) as an integer, one instruction */
__device__ __forceinline__ int floor_int(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

/* a * b + c for 24-bit signed a and b, one instruction */
__device__ __forceinline__ int mad_i24(int a, int b, int c)
{
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

/* the bilinear fraction: x - floor(x), capped at the largest float below 1
 * (v_fract_f32) */
__device__ __forceinline__ float fract1(float x)
{
    return __builtin_amdgcn_fractf(x);
}

/* Source coordinate of tap (dx, dy) of column c: byte offset of the
 * top-left texel of its bilinear footprint in the padded pool and the
 * bilinear fractions. */
template <int kPath>
__device__ __forceinline__ void tap_coord(const Warp& w, const Column& c, int pitch8, float dy, uint32_t* off,
                                          float* fa, float* fb)
{
    /* the linear reciprocal 1 - ex dx - ey dy, one fused multiply-add */
    const float lin = fmaf(-w.ey, dy, c.omex);
    float inv;
    if (kPath == kPathLinear) {
        inv = lin;
    } else {
        /* the cubic 1 + u (u (1 - u) - 1) */
        const float u = fmaf(w.ey, dy, c.ex);
        const float cubic = fmaf(u, fmaf(u, 1.0f - u, -1.0f), 1.0f);
        inv = w.tiny ? lin : ((kPath != kPathAny || w.small) ? cubic : 1.0f / (1.0f + u));
    }
    float lx = fmaf(fmaf(w.gxy, dy, c.gx), inv, w.fx0);
    float ly = fmaf(fmaf(w.gyy, dy, c.gy), inv, w.fy0);
    if (kPath == kPathSmall || kPath == kPathAny) {
        lx = fminf(fmaxf(lx, w.lox), w.hix);
        ly = fminf(fmaxf(ly, w.loy), w.hiy);
    }
    *fa = fract1(lx);
    *fb = fract1(ly);
    /* (ix0 + floor(lx), iy0 + floor(ly)) lies in [-1, W] x [-1, H]: its
     * footprint entry is inside the padded image, whose border repeats the
     * edge texels; |floor(ly)| <= 8 and pitch8 < 2^23 fit the 24-bit
     * product; the sum is the 32-bit unsigned byte offset */
    *off = ((uint32_t)floor_int(lx) << 3) + (uint32_t)mad_i24(floor_int(ly), pitch8, (int)w.base8);
}

/* The padded pool as a buffer resource (gfx9 word 3): offsets are unsigned
 * 32-bit, the pool is below 2^29 entries of 8 bytes. */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pool_rsrc(const uint32_t* pool)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)pool, (short)0, (int)0xffffffffu, 0x00020000);
}

/* the bilinear coefficients of a footprint (hpm_half.h): entry (x, y) holds
 * c0 | c1 << 16 and c2 | c3 << 16 as FP16, one 8-byte load */
struct Footprint {
    uint32_t lo, hi;
};

__device__ __forceinline__ Footprint footprint(__amdgpu_buffer_rsrc_t pool, uint32_t off)
{
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 v = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(pool, off, 0, 0));
    Footprint f;
    f.lo = v.x;
    f.hi = v.y;
    return f;
}

typedef _Float16 hpm_h2 __attribute__((ext_vector_type(2)));

/* linear-filtered, clamp-addressed fetch (the reference's tex2D on a
 * cudaFilterModeLinear texture, HPM.cpp:599-617) at fractions (fa, fb) of a
 * footprint: (c0 + fa c1) + fb (c2 + fa c3) as hpm_bilinear_eval, each
 * step one fused multiply-add, the first two v_fma_mix_f32 on the FP16
 * coefficients */
__device__ __forceinline__ float bilerp(Footprint q, float fa, float fb)
{
    const hpm_h2 l = __builtin_bit_cast(hpm_h2, q.lo);
    const hpm_h2 h = __builtin_bit_cast(hpm_h2, q.hi);
    const float v0 = fmaf(fa, (float)l.y, (float)l.x);
    const float v = fmaf(fa, (float)h.y, (float)h.x);
    return fmaf(fb, v, v0);
}

/* Bilateral weights of the reference taps, (w, w * (r - c)) per tap and
 * lane of a 256-lane block: [tap][lane]. In LDS they leave the VGPRs to the
 * gathers in flight; every lane reads only its own entries. */
__shared__ float2 s_refw[HPM_NTAP][256];

/* reference-only part of ComputeBilateralNCC + ComputeBilateralWeight
 * (HPM.cu:318-369): the window weights go to s_refw, the sums stay here */
struct RefPatch {
    int lane;
    float c;
    float inv_w;
    float mr;
    float var_r;
};

__device__ __forceinline__ void make_ref_patch(const float* img, int W, int H, int px, int py, const hpm_params& P,
                                               RefPatch* rp)
{
    const int lane = (int)(threadIdx.x + blockDim.x * threadIdx.y);
    const float c = texel(img, W, H, px, py);
    const float two_ss = 2.0f * P.sigma_spatial * P.sigma_spatial;
    const float two_sc = 2.0f * P.sigma_color * P.sigma_color;
    /* the 36 reads are issued together */
    float tex[HPM_NTAP];
#pragma unroll
    for (int k = 0; k < HPM_NTAP; ++k) {
        tex[k] = texel(img, W, H, px + 2 * (k / 6) - 5, py + 2 * (k % 6) - 5);
    }
    float wsum = 0.0f, sr = 0.0f, srr = 0.0f;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            const int i = 2 * a - 5, j = 2 * b - 5;
            const int k = 6 * a + b;
            const float r = tex[k];
            const float fi = (float)i, fj = (float)j;
            const float sd = sqrtf(fi * fi + fj * fj);
            const float cd = fabsf(r - c);
            const float w = hpm_expf(-sd / two_ss - cd / two_sc);
            const float rr = r - c;
            const float wr = w * rr;
            s_refw[k][lane] = make_float2(w, wr);
            wsum = wsum + w;
            sr = sr + wr;
            srr = fmaf(wr, rr, srr);
        }
    }
    rp->lane = lane;
    rp->c = c;
    rp->inv_w = 1.0f / wsum;
    rp->mr = sr * rp->inv_w;
    rp->var_r = srr * rp->inv_w - rp->mr * rp->mr;
}

/* One column of the window (6 taps at x offset dx): coordinates, then the
 * 6 gathers. */
template <int kPath>
__device__ __forceinline__ void fetch_column(__amdgpu_buffer_rsrc_t pool, int pitch8, const Warp& w, float dx,
                                             Footprint* quad, float* fa, float* fb)
{
    uint32_t off[6];
    const Column c = column<kPath>(w, dx);
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        tap_coord<kPath>(w, c, pitch8, (float)(2 * b - 5), &off[b], &fa[b], &fb[b]);
    }
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        quad[b] = footprint(pool, off[b]);
    }
}

/* Adds the 6 taps of window column a (taps 6a .. 6a + 5) to the sums, in
 * tap order; tap 14 is the shift itself. */
__device__ __forceinline__ void accumulate_column(const RefPatch& rp, int a, float shift, const Footprint* quad,
                                                  const float* fa, const float* fb, float* ss, float* sss, float* srs)
{
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const float raw = (a == 2 && b == 2) ? shift : bilerp(quad[b], fa[b], fb[b]);
        const float s = raw - shift;
        const float2 wk = s_refw[6 * a + b][rp.lane];
        const float ws = wk.x * s;
        *ss = *ss + ws;
        *sss = fmaf(ws, s, *sss);
        *srs = fmaf(wk.y, s, *srs);
    }
}

/* Window sums of one (hypothesis, source view), software-pipelined by
 * column: the gathers of column a + 1 are in flight while column a is
 * accumulated. Tap k = 6 a + b sits at offset (2a - 5, 2b - 5). */
template <int kPath>
__device__ __forceinline__ void ncc_sums(__amdgpu_buffer_rsrc_t pool, int pitch8, const RefPatch& rp, const Warp& w,
                                         float* ss_out, float* sss_out, float* srs_out)
{
    float ss = 0.0f, sss = 0.0f, srs = 0.0f;
    Footprint cur[6], nxt[6];
    float fa[6], fb[6], nfa[6], nfb[6];
    /* source values relative to the tap nearest the window centre (tap 14,
     * offset (-1, -1)), sampled first: a constant patch sums to exactly 0 */
    uint32_t coff;
    float cfa, cfb;
    tap_coord<kPath>(w, column<kPath>(w, -1.0f), pitch8, -1.0f, &coff, &cfa, &cfb);
    const Footprint cq = footprint(pool, coff);
    fetch_column<kPath>(pool, pitch8, w, -5.0f, cur, fa, fb);
    const float shift = bilerp(cq, cfa, cfb);
    /* two columns per trip with the buffers alternating; the last two
     * columns follow the loop, so every prefetch is unconditional and the
     * waits stay on the older loads */
#pragma unroll 1
    for (int a = 0; a < 4; a += 2) {
        fetch_column<kPath>(pool, pitch8, w, (float)(2 * a - 3), nxt, nfa, nfb);
        accumulate_column(rp, a, shift, cur, fa, fb, &ss, &sss, &srs);
        fetch_column<kPath>(pool, pitch8, w, (float)(2 * a - 1), cur, fa, fb);
        accumulate_column(rp, a + 1, shift, nxt, nfa, nfb, &ss, &sss, &srs);
    }
    fetch_column<kPath>(pool, pitch8, w, 5.0f, nxt, nfa, nfb);
    accumulate_column(rp, 4, shift, cur, fa, fb, &ss, &sss, &srs);
    accumulate_column(rp, 5, shift, nxt, nfa, nfb, &ss, &sss, &srs);
    *ss_out = ss;
    *sss_out = sss;
    *srs_out = srs;
}

/* ComputeBilateralNCC (HPM.cu:325-397) for source view v; q from plane_q */
__device__ __forceinline__ float ncc(const hpm_dev_problem& pb, const RefPatch& rp, int px, int py, int v,
                                     const double q[3])
{
    const hpm_cview sv = cview(pb, v);
    Warp w;
    if (!make_warp(sv, q, px, py, &w)) {
        return 2.0f;
    }
    const int pitch8 = sv->pitch * 8;
    const __amdgpu_buffer_rsrc_t pool = pool_rsrc(pb.images_pad);
    float ss, sss, srs;
    if (__all(w.inner)) {
        ncc_sums<kPathLinear>(pool, pitch8, rp, w, &ss, &sss, &srs);
    } else if (__all(w.small)) {
        ncc_sums<kPathSmall>(pool, pitch8, rp, w, &ss, &sss, &srs);
    } else {
        ncc_sums<kPathAny>(pool, pitch8, rp, w, &ss, &sss, &srs);
    }
    const float ms = ss * rp.inv_w;
    const float var_s = sss * rp.inv_w - ms * ms;
    const float kMinVar = 1e-5f;
    if (rp.var_r < kMinVar || var_s < kMinVar) {
        return 2.0f;
    }
    const float cov = srs * rp.inv_w - rp.mr * ms;
    /* var_r, var_s >= 1e-5 and both below 255^2: the product and its root are
     * normal, the quotient needs no scaling */
    const float den = sqrt_rn(rp.var_r * var_s);
    return fmaxf(0.0f, fminf(2.0f, 1.0f - div_rn(cov, den)));
}

/* HPM.cu:102-133 */
__device__ __forceinline__ void point3(const hpm_camera& c, int px, int py, float depth, float X[3])
{
    X[0] = depth * ((float)px - c.K[2]) / c.K[0];
    X[1] = depth * ((float)py - c.K[5]) / c.K[4];
    X[2] = depth;
}

__device__ __forceinline__ float4 view_direction(const hpm_camera& c, int px, int py, float depth)
{
    float X[3];
    point3(c, px, py, depth, X);
    const float norm = sqrtf(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    return make_float4(X[0] / norm, X[1] / norm, X[2] / norm, 0.0f);
}

__device__ __forceinline__ float dist_to_origin(const hpm_camera& c, int px, int py, float depth, float4 n)
{
    float X[3];
    point3(c, px, py, depth, X);
    return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}

__device__ __forceinline__ float depth_from_plane(const hpm_camera& c, float4 pl, int px, int py)
{
    return -pl.w * c.K[0] /
           (((float)px - c.K[2]) * pl.x + (c.K[0] / c.K[4]) * ((float)py - c.K[5]) * pl.y + c.K[0] * pl.z);
}

/* NormalizeVec3 (HPM.cu:77-84) with 1/sqrtf */
__device__ __forceinline__ void normalize3(float4* v)
{
    const float n2 = v->x * v->x + v->y * v->y + v->z * v->z;
    const float inv = 1.0f / sqrtf(n2);
    v->x *= inv;
    v->y *= inv;
    v->z *= inv;
}

__device__ __forceinline__ float dot3(float4 a, float4 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

/* GenerateRandomNormal (HPM.cu:135-161) */
__device__ float4 random_normal(const hpm_camera& c, int px, int py, hpm_rng* rng, float depth)
{
    float q1 = 1.0f, q2 = 1.0f, s = 2.0f;
    while (s >= 1.0f) {
        q1 = 2.0f * hpm_rng_uniform(rng) - 1.0f;
        q2 = 2.0f * hpm_rng_uniform(rng) - 1.0f;
        s = q1 * q1 + q2 * q2;
    }
    const float sq = sqrtf(1.0f - s);
    float4 n = make_float4(2.0f * q1 * sq, 2.0f * q2 * sq, 1.0f - 2.0f * s, 0.0f);
    const float4 vd = view_direction(c, px, py, depth);
    const float d = n.x * vd.x + n.y * vd.y + n.z * vd.z;
    if (d > 0.0f) {
        n.x = -n.x;
        n.y = -n.y;
        n.z = -n.z;
    }
    normalize3(&n);
    return n;
}

/* GeneratePerturbedNormal (HPM.cu:163-198) */
__device__ float4 perturbed_normal(const hpm_camera& c, int px, int py, float4 normal, hpm_rng* rng,
                                   float perturbation)
{
    const float4 vd = view_direction(c, px, py, 1.0f);
    const float a1 = (hpm_rng_uniform(rng) - 0.5f) * perturbation;
    const float a2 = (hpm_rng_uniform(rng) - 0.5f) * perturbation;
    const float a3 = (hpm_rng_uniform(rng) - 0.5f) * perturbation;
    const float s1 = hpm_sinf(a1), s2 = hpm_sinf(a2), s3 = hpm_sinf(a3);
    const float c1 = hpm_cosf(a1), c2 = hpm_cosf(a2), c3 = hpm_cosf(a3);
    float R[9];
    R[0] = c2 * c3;
    R[1] = c3 * s1 * s2 - c1 * s3;
    R[2] = s1 * s3 + c1 * c3 * s2;
    R[3] = c2 * s3;
    R[4] = c1 * c3 + s1 * s2 * s3;
    R[5] = c1 * s2 * s3 - c3 * s1;
    R[6] = -s2;
    R[7] = c2 * s1;
    R[8] = c1 * c2;
    float4 np;
    np.x = R[0] * normal.x + R[1] * normal.y + R[2] * normal.z;
    np.y = R[3] * normal.x + R[4] * normal.y + R[5] * normal.z;
    np.z = R[6] * normal.x + R[7] * normal.y + R[8] * normal.z;
    np.w = normal.w;
    if (dot3(np, vd) >= 0.0f) {
        np = normal;
    }
    normalize3(&np);
    return np;
}

/* Get3DPointonWorld_cu / ProjectonCamera_cu (HPM.cu:445-481) */
__device__ __forceinline__ void point_world(float x, float y, float depth, const hpm_camera& c, float out[3])
{
    float p[3], t[3], C[3];
    p[0] = depth * (x - c.K[2]) / c.K[0];
    p[1] = depth * (y - c.K[5]) / c.K[4];
    p[2] = depth;
    t[0] = c.R[0] * p[0] + c.R[3] * p[1] + c.R[6] * p[2];
    t[1] = c.R[1] * p[0] + c.R[4] * p[1] + c.R[7] * p[2];
    t[2] = c.R[2] * p[0] + c.R[5] * p[1] + c.R[8] * p[2];
    C[0] = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
    C[1] = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
    C[2] = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
    out[0] = t[0] + C[0];
    out[1] = t[1] + C[1];
    out[2] = t[2] + C[2];
}

__device__ __forceinline__ void project(const float X[3], const hpm_camera& c, float* u, float* v, float* d)
{
    float t[3];
    t[0] = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2] + c.t[0];
    t[1] = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2] + c.t[1];
    t[2] = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2] + c.t[2];
    *d = c.K[6] * t[0] + c.K[7] * t[1] + c.K[8] * t[2];
    *u = (c.K[0] * t[0] + c.K[1] * t[1] + c.K[2] * t[2]) / *d;
    *v = (c.K[3] * t[0] + c.K[4] * t[1] + c.K[5] * t[2]) / *d;
}

/* ComputeGeomConsistencyCost (HPM.cu:483-508; ACMM.cu:494-519 caps it at
 * 3), in two steps so that several depth reads can be in flight: the
 * projection into source view v and the read of its depth there (the
 * clamped index stays inside the map for any plane), then the
 * reprojection and the capped distance */
struct GeomSample {
    float sx, sy, src_depth;
};

__device__ __forceinline__ GeomSample geom_fetch(const hpm_dev_problem& pb, int v, float4 plane, int px, int py)
{
    const hpm_camera rc = camera(pb, 0);
    const hpm_camera sc = camera(pb, v);
    const float depth = depth_from_plane(rc, plane, px, py);
    float fw[3];
    point_world((float)px, (float)py, depth, rc, fw);
    GeomSample g;
    float sd;
    project(fw, sc, &g.sx, &g.sy, &sd);
    const float cxf = fminf(fmaxf(g.sx, -1.0f), (float)sc.width);
    const float cyf = fminf(fmaxf(g.sy, -1.0f), (float)sc.height);
    const int ix = clampi((int)cxf, 0, sc.width - 1);
    const int iy = clampi((int)cyf, 0, sc.height - 1);
    g.src_depth = pb.depths[cview(pb, v)->depth_off + (int64_t)iy * sc.width + ix];
    return g;
}

__device__ __forceinline__ float geom_finish(const hpm_dev_problem& pb, int v, GeomSample g, int px, int py)
{
    const float max_cost = pb.params.acmm ? 3.0f : 5.0f;
    if (g.src_depth == 0.0f) {
        return max_cost;
    }
    const hpm_camera rc = camera(pb, 0);
    const hpm_camera sc = camera(pb, v);
    float bw[3];
    point_world(g.sx, g.sy, g.src_depth, sc, bw);
    float bx, by, bd;
    project(bw, rc, &bx, &by, &bd);
    const float dx = (float)px - bx;
    const float dy = (float)py - by;
    return fminf(max_cost, sqrtf(dx * dx + dy * dy));
}

/* the empty asm keeps the first element a value: a select between two
 * loads of the array would be folded into one load from a selected
 * address, and the array (indexed at run time) would stay in memory */
__device__ __forceinline__ int find_min_idx8(const float* c)
{
    float m = c[0];
    asm volatile("" : "+v"(m));
    int idx = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        if (c[i] <= m) {
            m = c[i];
            idx = i;
        }
    }
    return idx;
}

__device__ __forceinline__ int find_max_idx8(const float* c)
{
    float m = c[0];
    asm volatile("" : "+v"(m));
    int idx = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        if (c[i] >= m) {
            m = c[i];
            idx = i;
        }
    }
    return idx;
}

/* ComputeMultiViewCostVector (HPM.cu:438-443); views with a zero weight
 * are skipped when weights is given (their terms are never read) */
__device__ void cost_vector(const hpm_dev_problem& pb, const RefPatch& rp, int px, int py, float4 plane, float* out,
                            const float* weights, uint32_t* nev)
{
    double q[3];
    plane_q(pb, plane, q);
    for (int v = 1; v <= pb.nsrc; ++v) {
        if (weights != nullptr && !(weights[v - 1] > 0.0f)) {
            out[v - 1] = 0.0f;
            continue;
        }
        out[v - 1] = ncc(pb, rp, px, py, v, q);
        ++*nev;
    }
}

/* ComputeMultiViewInitialCostandSelectedViews (HPM.cu:399-436) */
/* The initial cost and view mask of a hypothesis; with vc (the pixel's
 * cost-cache entry, views pb.W * pb.H apart) its per-view costs are kept. */
__device__ float initial_cost(const hpm_dev_problem& pb, const RefPatch& rp, int px, int py, float4 plane,
                              uint32_t* sel, uint32_t* nev, float* vc = nullptr)
{
    const size_t np = (size_t)pb.W * pb.H;
    const int nsrc = pb.nsrc;
    const int kparam = pb.params.top_k;
    float cv[HPM_MAX_SOURCES], cvc[HPM_MAX_SOURCES];
    /* for top_k <= 8 the eight smallest costs in ascending order, kept in
     * registers by a compare-and-swap insertion (an equal cost goes after
     * the ones already kept, as in the insertion sort) */
    const bool small_k = kparam <= 8;
    float best[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        best[j] = 3.0f; /* above every cost */
    }
    int count = 0, valid = 0;
    double q[3];
    plane_q(pb, plane, q);
    for (int v = 1; v <= nsrc; ++v) {
        const float c = ncc(pb, rp, px, py, v, q);
        ++*nev;
        if (vc) {
            vc[(size_t)(v - 1) * np] = c;
        }
        cvc[v - 1] = c;
        ++count;
        if (c < 2.0f) {
            ++valid;
        }
        if (small_k) {
            float x = c;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const bool lt = x < best[j];
                const float keep = lt ? x : best[j];
                x = lt ? best[j] : x;
                best[j] = keep;
            }
        } else {
            cv[v - 1] = c;
        }
    }
    uint32_t m = 0u;
    const int top_k = min(valid, kparam);
    float result = 2.0f;
    if (small_k) {
        if (top_k > 0) {
            float cost = 0.0f;
            float thr = best[0];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < top_k) {
                    cost += best[i];
                    thr = best[i];
                }
            }
            /* the views at or below the k-th cost, eight costs read at once */
#pragma unroll 1
            for (int v0 = 0; v0 < nsrc; v0 += 8) {
                float c[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    c[k] = cvc[min(v0 + k, nsrc - 1)];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (v0 + k < nsrc && c[k] <= thr) {
                        m |= (1u << (v0 + k));
                    }
                }
            }
            result = cost / (float)top_k;
        }
        *sel = m;
        return result;
    }
    /* sort_small (HPM.cu:3-12) */
    for (int i = 1; i < count; i++) {
        const float tmp = cv[i];
        int j;
        for (j = i; j >= 1 && tmp < cv[j - 1]; j--) {
            cv[j] = cv[j - 1];
        }
        cv[j] = tmp;
    }
    if (top_k > 0) {
        float cost = 0.0f;
        for (int i = 0; i < top_k; ++i) {
            cost += cv[i];
        }
        const float thr = cv[top_k - 1];
        for (int i = 0; i < nsrc; ++i) {
            if (cvc[i] <= thr) {
                m |= (1u << i);
            }
        }
        result = cost / (float)top_k;
    }
    *sel = m;
    return result;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
    }
    return v;
}

__device__ __forceinline__ void count_evals(const hpm_dev_problem& pb, uint32_t nev)
{
    if (pb.n_eval == nullptr) {
        return;
    }
    const unsigned long long s = wave_sum((unsigned long long)nev);
    if ((threadIdx.x + blockDim.x * threadIdx.y) % 64 == 0 && s != 0ull) {
        atomicAdd(pb.n_eval, s);
    }
}

/* JBU_cu (HPM.cu:1984-2052) for one output pixel: SpatialGauss /
 * RangeGauss (HPM.cu:1971-1981) in FP32, NormalizeVec3 with 1/sqrtf */
__device__ __forceinline__ void jbu_pixel(const float* guide, int w, int h, const float* src_nd, int sw, int sh,
                                          int scale_i, int px, int py, float* depth_out, float4* normal_out)
{
    const float scale = (float)(1.0 * sw / w);
    const float sigmad = 0.50f, sigmar = 25.5f;
    const int win = scale_i * scale_i + 1;
    const int nn = win / 2;
    const float o_y = (float)py * scale;
    const float o_x = (float)px * scale;
    const float ref_pix = texel(guide, w, h, px, py);
    float nf = 0.0f, ctot = 0.0f;
    float4 ntot = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int j = -nn; j <= nn; ++j) {
        int r_y = (int)(o_y + (float)j);
        r_y = (r_y > 0 ? (r_y < sh ? r_y : sh - 1) : 0);
        const int r_ys = py + j;
        for (int i = -nn; i <= nn; ++i) {
            int r_x = (int)(o_x + (float)i);
            r_x = (r_x > 0 ? (r_x < sw ? r_x : sw - 1) : 0);
            const int s_center = r_y * sw + r_x;
            float4 sn = ld4(src_nd, s_center);
            const float src_pix = sn.w;
            const float nb_pix = texel(guide, w, h, px + i, r_ys);
            const float ddx = o_x - (float)r_x, ddy = o_y - (float)r_y;
            const float dis = ddx * ddx + ddy * ddy;
            const float sg = hpm_expf(-dis / (2 * sigmad * sigmad));
            const float xp = fabsf(ref_pix - nb_pix);
            const float rg = hpm_expf(-(xp * xp) / (2 * sigmar * sigmar));
            const float tg = sg * rg;
            nf += tg;
            ctot += src_pix * tg;
            sn.x = sn.x * tg;
            sn.y = sn.y * tg;
            sn.z = sn.z * tg;
            ntot.x = ntot.x + sn.x;
            ntot.y = ntot.y + sn.y;
            ntot.z = ntot.z + sn.z;
        }
    }
    ntot.x = ntot.x / nf;
    ntot.y = ntot.y / nf;
    ntot.z = ntot.z / nf;
    normalize3(&ntot);
    *depth_out = ctot / nf;
    *normal_out = make_float4(ntot.x, ntot.y, ntot.z, 0.0f);
}

/* RandomInitialization (HPM.cu:510-558; ACMM.cu:521-621 adds the hierarchy
 * branch) */
__device__ uint32_t init_pixel(const hpm_dev_problem& pb, int px, int py)
{
    const hpm_params& P = pb.params;
    const hpm_camera rc = camera(pb, 0);
    const int center = py * pb.W + px;
    uint32_t nev = 0;
    hpm_rng rng;
    hpm_rng_init(&rng, P.seed_lo, P.seed_hi, (uint32_t)center, hpm_stream_init(pb.call_index), pb.salt);
    RefPatch rp;
    make_ref_patch(pb.images + cview(pb, 0)->img_off, pb.W, pb.H, px, py, P, &rp);
    float4 plane;
    if (P.geom_consistency) {
        const float4 t = ld4(pb.hyp, center);
        plane.x = rc.R[0] * t.x + rc.R[1] * t.y + rc.R[2] * t.z;
        plane.y = rc.R[3] * t.x + rc.R[4] * t.y + rc.R[5] * t.z;
        plane.z = rc.R[6] * t.x + rc.R[7] * t.y + rc.R[8] * t.z;
        plane.w = t.w;
        const float depth = plane.w;
        plane.w = dist_to_origin(rc, px, py, depth, plane);
    } else if (P.hierarchy) {
        /* the upsample branch (ACMM.cu:532-601): the depth is the JBU of
         * the coarse depth (RunJBU, ACMM.cpp:804-853, written to and read
         * back from depths.dmb), the normal the JBU of the coarse normals in
         * the same window, normalised. The pre-cost is the cost of the
         * plane (0, 0, 0, depth): the reference reads the normal of a fresh
         * host array there (ACMM.cpp:711-717), taken as zeros. */
        float depth;
        float4 n;
        jbu_pixel(pb.images + cview(pb, 0)->img_off, pb.W, pb.H, pb.coarse, pb.coarse_w, pb.coarse_h,
                  pb.coarse_scale, px, py, &depth, &n);
        uint32_t m0;
        pb.pre_cost[center] = initial_cost(pb, rp, px, py, make_float4(0.0f, 0.0f, 0.0f, depth), &m0, &nev);
        plane.x = rc.R[0] * n.x + rc.R[1] * n.y + rc.R[2] * n.z;
        plane.y = rc.R[3] * n.x + rc.R[4] * n.y + rc.R[5] * n.z;
        plane.z = rc.R[6] * n.x + rc.R[7] * n.y + rc.R[8] * n.z;
        plane.w = dist_to_origin(rc, px, py, depth, plane);
    } else if (P.planar_prior) {
        if (pb.prior_mask[center] > 0u && pb.cost[center] >= 0.1f) {
            const float perturbation = 0.02f;
            const float4 pl = ld4(pb.prior_planes, center);
            float dp = pl.w;
            const float dmin_p = (1 - 3 * perturbation) * dp;
            const float dmax_p = (1 + 3 * perturbation) * dp;
            dp = hpm_rng_uniform(&rng) * (dmax_p - dmin_p) + dmin_p;
            plane = perturbed_normal(rc, px, py, pl, &rng, (float)((double)(3 * perturbation) * kPi));
            plane.w = dp;
        } else {
            plane = ld4(pb.hyp, center);
            const float depth = plane.w;
            plane.w = dist_to_origin(rc, px, py, depth, plane);
        }
    } else {
        const float depth = hpm_rng_uniform(&rng) * (P.depth_max - P.depth_min) + P.depth_min;
        plane = random_normal(rc, px, py, &rng, depth);
        plane.w = dist_to_origin(rc, px, py, depth, plane);
    }
    st4(pb.hyp, center, plane);
    uint32_t m;
    pb.cost[center] = initial_cost(pb, rp, px, py, plane, &m, &nev, pb.vcost ? pb.vcost + center : nullptr);
    pb.sel[center] = m;
    return nev;
}

struct PriorConsts {
    float gamma, tdss, tass, beta;
};

__device__ __forceinline__ PriorConsts prior_consts(const hpm_params& P)
{
    PriorConsts k;
    k.gamma = 0.5f;
    const float ds = (P.depth_max - P.depth_min) / 64.0f;
    k.tdss = 2 * ds * ds;
    const float as = (float)(kPi * (double)(5.0f / 180.0f));
    k.tass = 2 * as * as;
    k.beta = 0.18f;
    return k;
}

__device__ __forceinline__ float prior_term(const PriorConsts& k, float depth_diff, float4 prior_plane, float4 plane)
{
    const float ac = dot3(prior_plane, plane);
    const float ad = hpm_acosf(ac);
    return k.gamma + hpm_expf(-depth_diff * depth_diff / k.tdss) * hpm_expf(-ad * ad / k.tass);
}

/* CheckerboardPropagation (HPM.cu:1222-1725) + PlaneHypothesisRefinement
 * (HPM.cu:1143-1220).
 *
 * The reference scores hypotheses phase by phase (eight neighbours, the
 * NESP extension, the current hypothesis, five refinement candidates); here
 * the wave runs the phases in step, so the decision code between them (view
 * selection, adoption, candidate generation, the refinement decisions) runs
 * once per wave rather than once for every group of lanes that reaches it.
 * Per-view results are stored and summed in view order, so every value and
 * every sum order is the sequential one. */
__device__ __forceinline__ int lane_rank(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* the reference sums of lane `owner` of wave wv; its window weights are
 * read from LDS at its lane index */
__device__ __forceinline__ RefPatch shfl_ref_patch(const RefPatch& rp, int wv, int owner)
{
    RefPatch o;
    o.lane = wv * 64 + owner;
    o.c = __shfl(rp.c, owner);
    o.inv_w = __shfl(rp.inv_w, owner);
    o.mr = __shfl(rp.mr, owner);
    o.var_r = __shfl(rp.var_r, owner);
    return o;
}

/* the number of the 15 view-selection draws per source view, 4 bits per
 * view (views 0-15 in lo, 16-31 in hi): registers rather than an array */
struct DrawCounts {
    uint64_t lo, hi;
};

__device__ __forceinline__ float draws_of(const DrawCounts& d, int v)
{
    const uint64_t w = v < 16 ? d.lo : d.hi;
    return (float)(uint32_t)((w >> (4 * (v & 15))) & 15u);
}

__device__ __forceinline__ void set_draws(DrawCounts* d, int v, float n)
{
    const uint64_t b = (uint64_t)(uint32_t)n << (4 * (v & 15));
    if (v < 16) {
        d->lo |= b;
    } else {
        d->hi |= b;
    }
}

/* the candidates c of a view whose tasks [pv + c n, pv + (c + 1) n) meet
 * the trip [t0, t0 + 64): cl .. ch */
template <int C>
__device__ __forceinline__ void cand_range(int pv, int n, int t0, int* cl, int* ch)
{
    int lo = 0, hi = 0;
#pragma unroll
    for (int c = 1; c < C; ++c) {
        if (pv + c * n <= t0) {
            lo = c;
        }
        if (pv + c * n < t0 + 64) {
            hi = c;
        }
    }
    *cl = lo;
    *ch = hi;
}

/* per wave of a 256-lane block: the lanes that selected the view being
 * refined, in lane order */
__shared__ int s_own[4][64];

/* per wave of a 256-lane block: the tasks of the current packed trip,
 * owner lane | view << 8 | candidate << 16 */
__shared__ int s_task[4][64];

/* Packed evaluation in view-major order. For each source view v, in
 * ascending order, the lanes with bit v of vmask each score C candidate
 * planes on v; task t runs over (v, c, r): candidate c of the r-th such lane
 * in lane order. Every trip of the wave takes the next 64 tasks: lane j
 * scores task t0 + j for its owner, whose pixel, reference sums and
 * candidate come by lane permutes (the window weights are in LDS at the
 * owner's index), so only the last trip runs partly empty. take(v, c, value)
 * then runs on the owner for each of its tasks in the trip, views
 * ascending: per candidate, every owner receives its views in ascending
 * order. Every lane of the wave calls this. */
template <int C, typename PlaneOf, typename Score, typename Take>
__device__ __forceinline__ void packed_views(int nsrc, const RefPatch& rp, int px, int py, uint32_t vmask,
                                             PlaneOf&& plane_of, Score&& score, Take&& take)
{
    const int wv = (int)threadIdx.y, ln = (int)threadIdx.x; /* blockDim = (64, 4) */
    int total = 0;
#pragma unroll 1
    for (int v = 0; v < nsrc; ++v) {
        total += C * __popcll(__ballot((vmask >> v) & 1u));
    }
    const int pxy = px | (py << 16);
    int va = 0, pa = 0; /* the first view with tasks at or after t0, and its first task */
#pragma unroll 1
    for (int t0 = 0; t0 < total; t0 += 64) {
#pragma unroll 1
        for (;;) {
            const int na = C * __popcll(__ballot((vmask >> va) & 1u));
            if (pa + na > t0) {
                break;
            }
            pa += na;
            ++va;
        }
        /* the trip's task table and the candidates it covers */
        int c_lo = C - 1, c_hi = 0;
        int vb = va;
#pragma unroll 1
        for (int pv = pa; vb < nsrc && pv < t0 + 64; ++vb) {
            const bool mine = ((vmask >> vb) & 1u) != 0u;
            const uint64_t m = __ballot(mine);
            const int n = __popcll(m);
            if (n > 0) {
                int cl, ch;
                cand_range<C>(pv, n, t0, &cl, &ch);
                c_lo = min(c_lo, cl);
                c_hi = max(c_hi, ch);
                const int slot = pv + lane_rank(m) - t0;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int i = slot + c * n;
                    if (mine && c >= cl && c <= ch && i >= 0 && i < 64) {
                        s_task[wv][i] = ln | (vb << 8) | (c << 16);
                    }
                }
            }
            pv += C * n;
        }
        __builtin_amdgcn_wave_barrier();
        const bool act = t0 + ln < total;
        const int e = act ? s_task[wv][ln] : ln;
        const int owner = e & 0xff, v = (e >> 8) & 0xff, c = e >> 16;
        const int opxy = __shfl(pxy, owner);
        const RefPatch orp = shfl_ref_patch(rp, wv, owner);
        float4 plane = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            if (cc >= c_lo && cc <= c_hi) {
                const float4 mp = plane_of(cc);
                const float4 pc =
                    make_float4(__shfl(mp.x, owner), __shfl(mp.y, owner), __shfl(mp.z, owner), __shfl(mp.w, owner));
                if (c == cc) {
                    plane = pc;
                }
            }
        }
        float term = 0.0f;
        if (act) {
            term = score(orp, opxy & 0xffff, opxy >> 16, v, plane);
        }
        /* each owner takes back its terms */
        int pv = pa;
#pragma unroll 1
        for (int vv = va; vv < vb; ++vv) {
            const bool mine = ((vmask >> vv) & 1u) != 0u;
            const uint64_t m = __ballot(mine);
            const int n = __popcll(m);
            if (n > 0) {
                int cl, ch;
                cand_range<C>(pv, n, t0, &cl, &ch);
                const int slot = pv + lane_rank(m) - t0;
#pragma unroll
                for (int cc = 0; cc < C; ++cc) {
                    if (cc >= cl && cc <= ch) {
                        const int src = slot + cc * n; /* the lane that scored (vv, cc) for this lane */
                        const bool got = mine && src >= 0 && src < 64;
                        const float val = __shfl(term, got ? src : ln);
                        if (got) {
                            take(vv, cc, val);
                        }
                    }
                }
            }
            pv += C * n;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

/* a[i] for an index known only at run time, by selects: the array is
 * never indexed dynamically, so it stays in registers. The first element
 * passes through an empty asm statement: a select between two loads would
 * otherwise be folded into one load from a selected address, which indexes
 * the array after all. */
template <typename T>
__device__ __forceinline__ T pick8(const T (&a)[8], int i)
{
    T r = a[0];
    asm volatile("" : "+v"(r));
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        r = i == k ? a[k] : r;
    }
    return r;
}

/* f[i] for an index known only at run time, from the flags' bit mask */
__device__ __forceinline__ bool pick8(const bool (&f)[8], int i)
{
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        m |= f[k] ? 1u << k : 0u;
    }
    return ((m >> i) & 1u) != 0u;
}

/* one source view's eight direction costs */
struct CostRow {
    float4 lo, hi;
};

/* f(i, load(i)) for i = 0 .. n - 1 in order (n >= 1): the loads of rows
 * i + 1 and i + 2 are in flight while row i is used, in two buffers, so no
 * copy between trips waits on a load. load(i) must be valid for i < n. */
template <typename Load, typename F>
__device__ __forceinline__ void pipelined(int n, Load&& load, F&& f)
{
    auto a = load(0);
#pragma unroll 1
    for (int i = 0; i < n; i += 2) {
        const auto b = load(min(i + 1, n - 1));
        f(i, a);
        if (i + 1 < n) {
            a = load(min(i + 2, n - 1));
            f(i + 1, b);
        }
    }
}

/* Every lane of the wave calls this (the refinement packs evaluations of
 * all lanes into the wave's trips); lanes without a pixel (valid false)
 * only help. */
__device__ __forceinline__ uint32_t propagate_pixel(const hpm_dev_problem& pb, int px, int py, int iter, int colour,
                                                    bool valid)
{
    if (!valid) { /* any pixel inside the image: its reads stay in bounds */
        px = 0;
        py = 0;
    }
    const hpm_params& P = pb.params;
    const hpm_camera rc = camera(pb, 0);
    const int W = pb.W, H = pb.H, nsrc = pb.nsrc;
    const float* cost = pb.cost;
    const int center = py * W + px;
    uint32_t nev = 0;
    hpm_rng rng;
    hpm_rng_init(&rng, P.seed_lo, P.seed_hi, (uint32_t)center, hpm_stream_update(pb.call_index, iter, colour),
                 pb.salt);
    RefPatch rp;
    make_ref_patch(pb.images + cview(pb, 0)->img_off, W, H, px, py, P, &rp);

    int pos[8];
    bool flag[8];
    pos[0] = center - 5 - 6 * W;  /* left_up */
    pos[1] = center - W * 5;      /* up_far */
    pos[2] = center + 6 - 5 * W;  /* right_up */
    pos[3] = center + W * 5;      /* down_far */
    pos[4] = center + 5 + 6 * W;  /* right_down */
    pos[5] = center - 5;          /* left_far */
    pos[6] = center - 6 + 5 * W;  /* left_down */
    pos[7] = center + 5;          /* right_far */

    /* axis strips (HPM.cu:1251-1341): start sample, up to 4 more two apart,
     * the strip ending at the first sample outside the image; keeps the
     * first minimum, the first maximum for right_far. A strip leaves the
     * image for good once it leaves it, so the samples inside are a prefix:
     * all five are loaded at once (those outside read the pixel's own cost,
     * a distinct address, and are not compared) */
    flag[1] = py > 4;
    flag[3] = py < H - 5;
    flag[5] = px > 4;
    flag[7] = px < W - 5;
    {
        const int step[4] = {-2 * W, 2 * W, -2, 2};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int s = 2 * k + 1;
            if (!flag[s]) {
                continue;
            }
            float cv[5];
            bool in[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                in[i] = k == 0 ? py > 4 + 2 * i : (k == 1 ? py < H - 5 - 2 * i : (k == 2 ? px > 4 + 2 * i : px < W - 5 - 2 * i));
                cv[i] = cost[in[i] ? pos[s] + i * step[k] : center];
            }
            float cmin = cv[0];
            int cmp = pos[s];
#pragma unroll
            for (int i = 1; i < 5; ++i) {
                const bool better = k == 3 ? cmin < cv[i] : cv[i] < cmin; /* right_far keeps the maximum (HPM.cu:1330) */
                if (in[i] && better) {
                    cmin = cv[i];
                    cmp = pos[s] + i * step[k];
                }
            }
            pos[s] = cmp;
        }
    }
    /* diagonal zig-zags (HPM.cu:1344-1471): start offset, first step axis,
     * per-axis step signs */
    {
        const int sx0[4] = {-5, 6, 5, -6};
        const int sy0[4] = {-6, -5, 6, 5};
        const int first_x[4] = {1, 0, 1, 0}; /* even steps move x (1) or y (0) */
        const int dxs[4] = {-2, 2, 2, -2};
        const int dys[4] = {-2, -2, 2, 2};
        const int slot[4] = {0, 2, 4, 6};
        flag[0] = px > 4 && py > 5;
        flag[2] = px < W - 6 && py > 4;
        flag[4] = px < W - 5 && py < H - 6;
        flag[6] = px > 5 && py < H - 5;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int s = slot[d];
            if (!flag[s]) {
                continue;
            }
            /* the path moves away from the image on both axes, so once a
             * step leaves the image every later one does: breaking there
             * (HPM.cu:1344-1460) and skipping the outside steps (left_down,
             * which never breaks, HPM.cu:1461-1470) select the same; the
             * seven samples are loaded at once */
            const float c0 = cost[pos[s]];
            int pts[7];
            bool in[7];
            float cv[7];
            int qx = px + sx0[d], qy = py + sy0[d];
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const bool move_x = ((i % 2 == 0) == (first_x[d] == 1));
                if (move_x) {
                    qx += dxs[d];
                } else {
                    qy += dys[d];
                }
                in[i] = qx >= 0 && qx < W && qy < H && qy >= 0;
                pts[i] = qy * W + qx;
                cv[i] = cost[in[i] ? pts[i] : center];
            }
            float cmin = c0;
            int cmp = pos[s];
#pragma unroll
            for (int i = 0; i < 7; i++) {
                if (in[i] && cv[i] < cmin) {
                    cmin = cv[i];
                    cmp = pts[i];
                }
            }
            pos[s] = cmp;
        }
    }

    /* cost_array[8][32] = { 2.0f } (HPM.cu:1244), held view-major here:
     * ca[v][d] is direction d on source view v, so a view's eight costs are
     * two 16-byte loads in the view selection and the candidate scoring */
    float ca[HPM_MAX_SOURCES][8] __attribute__((aligned(16)));
    for (int v = 0; v < nsrc; ++v) {
        reinterpret_cast<float4*>(ca[v])[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        reinterpret_cast<float4*>(ca[v])[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    ca[0][0] = 2.0f;
    const uint32_t all_views = nsrc >= 32 ? 0xffffffffu : ((1u << nsrc) - 1u);
    /* a view's eight costs */
    auto cost_row = [&](int v) {
        CostRow r;
        r.lo = reinterpret_cast<const float4*>(ca[v])[0];
        r.hi = reinterpret_cast<const float4*>(ca[v])[1];
        return r;
    };

    /* state handed between the phases */
    DrawCounts vw = {0ull, 0ull}; /* draws per source view */
    uint32_t sel_views = 0u; /* views with vw > 0 */
    uint32_t tsel = 0u;
    float wn = 0.0f;
    float fc[8];
    int min_idx = 0;
    const float4 cur = ld4(pb.hyp, center);
    /* the cost cache: when view 0's entry is not NaN, every view's entry is
     * the photometric cost of cur on that view (computed when cur was set,
     * by the same function), and the rescore takes it from there */
    const size_t np = (size_t)W * H;
    float* const vc = pb.vcost ? pb.vcost + center : nullptr;
    const bool known = valid && vc != nullptr && vc[0] == vc[0];
    /* where the stored hypothesis comes from: -1 cur, 0-7 that neighbour
     * direction (its costs on every view are in ca), 8 a refinement
     * candidate */
    int plane_src = -1;
    float cost_now = 0.0f;
    float cost_center = 0.0f;
    float4 plane_center = cur;
    uint32_t sel_center = 0u;
    float depth_now = 0.0f;
    float restricted_cost = 0.0f;
    const PriorConsts pk = prior_consts(P);
    const bool prior_on = P.planar_prior != 0;
    const bool prior_pix = prior_on && pb.prior_mask[center] > 0u;
    float4 prior_plane = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float depth_prior = 0.0f;
    float4 cand[5];
    float cand_d[5];

    if (!valid) {
        for (int d = 0; d < 8; ++d) {
            flag[d] = false;
        }
    }
    double q[3];
    /* the eight neighbour hypotheses on every view: the same work for every
     * lane away from the image border, so the wave runs it in lockstep; the
     * next direction's hypothesis is read while this one is scored */
    float4 next_plane = ld4(pb.hyp, flag[0] ? pos[0] : center);
#pragma unroll 1
    for (int d = 0; d < 8; ++d) {
        const float4 plane_d = next_plane;
        if (d + 1 < 8) {
            next_plane = ld4(pb.hyp, pick8(flag, d + 1) ? pick8(pos, d + 1) : center);
        }
        if (pick8(flag, d)) {
            plane_q(pb, plane_d, q);
#pragma unroll 1
            for (int v = 0; v < nsrc; ++v) {
                ca[v][d] = ncc(pb, rp, px, py, v + 1, q);
            }
            nev += (uint32_t)nsrc;
        }
    }

    /* The rest: up to three NESP extension rounds, joint view selection,
     * the rescore of the current hypothesis, adoption and five refinement
     * candidates. The phases run in step across the wave, so each decision
     * runs once per wave; the extension, rescore and refinement evaluations
     * of all lanes are packed into full trips (packed_views). Values, sum
     * orders and decisions per lane are those of the sequential order. */

    /* NESP extension, orientation 0 only (HPM.cu:1473-1543, 780-843): up to
     * 3 rounds, each judged on direction 0's costs */
    bool ext_on = flag[0];
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        bool extend = false;
        if (ext_on) {
            /* JudgeExtend (HPM.cu:1036-1059) */
            const float good = pb.good[iter][k];
            int good_sum = 0, bad_sum = 0;
            pipelined(
                nsrc, [&](int v) { return ca[v][0]; },
                [&](int, float c) {
                    good_sum += c < good ? 1 : 0;
                    bad_sum += c > 1.2f ? 1 : 0; /* float bad_threshold (HPM.cu:1042) */
                });
            if (good_sum >= 1 && bad_sum <= 2) {
                ext_on = false;
            } else {
                int lu = center - 5 - 6 * W - 4 * 2 * k - 4 * 2 * W * k;
                if (px > 4 + 8 * k && py > 5 + 8 * k) {
                    bool symbol = true;
                    /* a left-up path: once outside, always outside, so the
                     * break (HPM.cu:797-815) selects as skipping; the samples
                     * are loaded at once */
                    const float c0 = cost[lu];
                    const float cprev = cost[pos[0]];
                    int pts[7];
                    bool in[7];
                    float cv[7];
                    int qx = px - 5 - 8 * k, qy = py - 6 - 8 * k;
#pragma unroll
                    for (int i = 0; i < 7; i++) {
                        if (i % 2 == 0) {
                            qx -= 2;
                        } else {
                            qy -= 2;
                        }
                        in[i] = qx >= 0 && qx < W && qy < H && qy >= 0;
                        pts[i] = qy * W + qx;
                        cv[i] = cost[in[i] ? pts[i] : center];
                    }
                    float cmin = c0;
                    int cmp = lu;
#pragma unroll
                    for (int i = 0; i < 7; i++) {
                        if (in[i] && cv[i] < cmin) {
                            cmin = cv[i];
                            cmp = pts[i];
                        }
                    }
                    if (cmin > cprev) {
                        cmp = pos[0];
                        symbol = false;
                    }
                    lu = cmp;
                    extend = symbol && lu != pos[0];
                } else {
                    lu = pos[0];
                }
                pos[0] = lu;
            }
        }
        if (__any(extend)) {
            if (extend) {
                nev += (uint32_t)nsrc;
            }
            const float4 ext_plane = extend ? ld4(pb.hyp, pos[0]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            packed_views<1>(
                nsrc, rp, px, py, extend ? all_views : 0u, [&](int) { return ext_plane; },
                [&](const RefPatch& orp, int opx, int opy, int v, float4 plane) {
                    double tq[3];
                    plane_q(pb, plane, tq);
                    return ncc(pb, orp, opx, opy, v + 1, tq);
                },
                [&](int v, int, float val) { ca[v][0] = val; });
        }
    }

    {
        /* joint view selection (HPM.cu:1548-1613) */
        float probs[HPM_MAX_SOURCES];
        float psum = 0.0f; /* their sum in view order */
        {
            /* view priors from the four axis neighbours' selections, summed
             * in neighbour order per view */
            const int nb[4] = {center - W * 5, center + W * 5, center - 5, center + 5};
            bool nbf[4];
            uint32_t nbm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                nbf[k] = flag[2 * k + 1];
                nbm[k] = nbf[k] ? pb.sel[nb[k]] : 0u;
            }
            const float thr = pb.thr[iter];
            const float fallback = pb.fallback[iter];
            /* one view's probability from its eight costs (c in [0, 2]:
             * c^2 is 0 or at least 2^-48, so div_rn is the IEEE quotient;
             * every term is computed and the ones below thr summed in order) */
            auto view_prob = [&](int i, float4 c_lo, float4 c_hi) {
                float vprior = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (nbf[k]) {
                        vprior += ((nbm[k] >> i) & 1u) == 1u ? 0.9f : 0.1f;
                    }
                }
                float count = 0.0f;
                int count_false = 0;
                float tmpw = 0.0f;
                const float cs[8] = {c_lo.x, c_lo.y, c_lo.z, c_lo.w, c_hi.x, c_hi.y, c_hi.z, c_hi.w};
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float c = cs[j];
                    const float e = hpm_expf(div_rn(c * c, -0.18f));
                    const bool in = c < thr;
                    tmpw = in ? tmpw + e : tmpw;
                    count = in ? count + 1.0f : count;
                    count_false += c > 1.2f ? 1 : 0;
                }
                float pr = 0.0f;
                if (count > 2 && count_false < 3) {
                    pr = div_rn(tmpw, count);
                } else if (count_false < 3) {
                    pr = fallback;
                }
                probs[i] = pr * vprior;
                psum += probs[i];
            };
            pipelined(nsrc, cost_row, [&](int i, CostRow r) { view_prob(i, r.lo, r.hi); });
        }
        {
            /* TransformPDFToCDF (HPM.cu:86-100) and the 15 draws: every draw
             * goes to the first view whose CDF value exceeds it; the draws
             * are made first and assigned in one pass over the views */
            const float inv = 1.0f / psum;
            float u[15];
#pragma unroll
            for (int k = 0; k < 15; ++k) {
                u[k] = hpm_rng_uniform(&rng) - FLT_EPSILON;
            }
            uint32_t open = 0x7fffu; /* draws not yet assigned */
            float cum = 0.0f;
            pipelined(nsrc, [&](int i) { return probs[i]; }, [&](int i, float p) {
                cum += p * inv;
                float n = 0.0f;
#pragma unroll
                for (int k = 0; k < 15; ++k) {
                    if (((open >> k) & 1u) != 0u && cum > u[k]) {
                        open &= ~(1u << k);
                        n += 1.0f;
                    }
                }
                set_draws(&vw, i, n);
            });
        }
        for (int i = 0; i < nsrc; ++i) {
            if (draws_of(vw, i) > 0) {
                tsel |= (1u << i);
                wn += draws_of(vw, i);
            }
        }
        sel_views = valid ? tsel : 0u;

        /* candidate scoring (HPM.cu:1615-1635); views outer, so the eight
         * directions' terms of a view load together, each direction's sum
         * still in view order */
        {
            float a[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                a[i] = 0.0f;
            }
            if (P.geom_consistency) {
                float4 hp[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    hp[i] = flag[i] ? ld4(pb.hyp, pos[i]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                pipelined(nsrc, cost_row, [&](int j, CostRow r) {
                    if ((tsel >> j) & 1u) {
                        const float w = draws_of(vw, j);
                        const float cs[8] = {r.lo.x, r.lo.y, r.lo.z, r.lo.w, r.hi.x, r.hi.y, r.hi.z, r.hi.w};
                        GeomSample gs[8];
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            gs[i] = geom_fetch(pb, j + 1, hp[i], px, py);
                        }
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const float g = flag[i] ? geom_finish(pb, j + 1, gs[i], px, py) : 5.0f;
                            a[i] += w * (cs[i] + 0.1f * g);
                        }
                    }
                });
            } else {
                pipelined(nsrc, cost_row, [&](int j, CostRow r) {
                    if ((tsel >> j) & 1u) {
                        const float w = draws_of(vw, j);
                        const float cs[8] = {r.lo.x, r.lo.y, r.lo.z, r.lo.w, r.hi.x, r.hi.y, r.hi.z, r.hi.w};
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            a[i] += w * cs[i];
                        }
                    }
                });
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                fc[i] = a[i] / wn;
            }
        }
        min_idx = find_min_idx8(fc);
    }

    /* weight of the geometric term in the re-score and the refinement:
     * 0.1 (HPM.cu:1640, 1194), 0.2 in ACMM (ACMM.cu:1596, 651) */
    const float gc_weight = P.acmm ? 0.2f : 0.1f;
    /* the term of one hypothesis on one selected view: the NCC cost, plus
     * the weighted geometric cost in geometric-consistency mode */
    auto selected_term = [&](const RefPatch& orp, int opx, int opy, int v, float4 plane) {
        GeomSample gs = {0.0f, 0.0f, 0.0f};
        if (P.geom_consistency) {
            gs = geom_fetch(pb, v + 1, plane, opx, opy); /* in flight during the NCC */
        }
        double tq[3];
        plane_q(pb, plane, tq);
        const float cv = ncc(pb, orp, opx, opy, v + 1, tq);
        return P.geom_consistency ? cv + gc_weight * geom_finish(pb, v + 1, gs, opx, opy) : cv;
    };

    /* re-score the current hypothesis on the selected views
     * (HPM.cu:1637-1650), packed over the wave's lanes and views; the views
     * arrive in ascending order, so the weighted sum accumulates in view
     * order */
    float rescore_sum = 0.0f;
    packed_views<1>(
        nsrc, rp, px, py, known ? 0u : sel_views, [&](int) { return cur; }, selected_term,
        [&](int v, int, float val) { rescore_sum += draws_of(vw, v) * val; });
    if (known) {
        /* the same terms from the cache, in the same view order; eight
         * views' entries are read at once */
#pragma unroll 1
        for (int v0 = 0; v0 < nsrc; v0 += 8) {
            float c[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                c[k] = vc[(size_t)min(v0 + k, nsrc - 1) * np];
            }
            GeomSample gs[8];
            if (P.geom_consistency) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    gs[k] = geom_fetch(pb, min(v0 + k, nsrc - 1) + 1, cur, px, py);
                }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int v = v0 + k;
                if (v < nsrc && ((sel_views >> v) & 1u)) {
                    const float term = P.geom_consistency ? c[k] + gc_weight * geom_finish(pb, v + 1, gs[k], px, py)
                                                          : c[k];
                    rescore_sum += draws_of(vw, v) * term;
                }
            }
        }
    }
    nev += (uint32_t)__popc(sel_views);

    {
        cost_now = rescore_sum / wn;
        cost_center = cost_now;
        plane_center = cur;
        sel_center = pb.sel[center];
        depth_now = depth_from_plane(rc, cur, px, py);

        /* adoption (HPM.cu:1652-1722) */
        if (prior_on) {
            prior_plane = ld4(pb.prior_planes, center);
            depth_prior = depth_from_plane(rc, prior_plane, px, py);
            if (prior_pix) {
                float rfc[8];
                float4 hps[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    hps[i] = ld4(pb.hyp, flag[i] ? pos[i] : center);
                }
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    rfc[i] = 0.0f;
                    if (flag[i]) {
                        const float4 hp = hps[i];
                        const float dn = depth_from_plane(rc, hp, px, py);
                        const float pr = prior_term(pk, dn - depth_prior, prior_plane, hp);
                        rfc[i] = hpm_expf(-fc[i] * fc[i] / pk.beta) * pr;
                    }
                }
                const int max_idx = find_max_idx8(rfc);
                const float dn = depth_from_plane(rc, cur, px, py);
                const float pr = prior_term(pk, dn - depth_prior, prior_plane, cur);
                const float rcn = hpm_expf(-cost_now * cost_now / pk.beta) * pr;
                if (pick8(flag, max_idx)) {
                    const float4 hp = ld4(pb.hyp, pick8(pos, max_idx));
                    const float db = depth_from_plane(rc, hp, px, py);
                    if (db >= P.depth_min && db <= P.depth_max && pick8(rfc, max_idx) > rcn) {
                        depth_now = db;
                        plane_center = hp;
                        plane_src = max_idx;
                        cost_center = pick8(fc, max_idx);
                        restricted_cost = pick8(rfc, max_idx);
                        sel_center = tsel;
                    }
                }
            } else if (pick8(flag, min_idx)) {
                const float4 hp = ld4(pb.hyp, pick8(pos, min_idx));
                const float db = depth_from_plane(rc, hp, px, py);
                if (db >= P.depth_min && db <= P.depth_max && pick8(fc, min_idx) < cost_now) {
                    depth_now = db;
                    plane_center = hp;
                    plane_src = min_idx;
                    cost_center = pick8(fc, min_idx);
                }
            }
        }
        if (!prior_on && pick8(flag, min_idx)) {
            const float4 hp = ld4(pb.hyp, pick8(pos, min_idx));
            const float db = depth_from_plane(rc, hp, px, py);
            if (db >= P.depth_min && db <= P.depth_max && pick8(fc, min_idx) < cost_now) {
                depth_now = db;
                plane_center = hp;
                plane_src = min_idx;
                cost_center = pick8(fc, min_idx);
                sel_center = tsel;
            }
            if (!P.acmm && cost_center != cost_center) {
                depth_now = db;
                plane_center = hp;
                plane_src = min_idx;
                cost_now = pick8(fc, min_idx);
                sel_center = tsel;
            }
        }

        /* PlaneHypothesisRefinement candidates (HPM.cu:1143-1190) */
        {
            const float perturbation = 0.02f;
            float depth_rand;
            float4 n_rand;
            const float ds = (P.depth_max - P.depth_min) / 64.0f;
            const float as = (float)(kPi * (double)(5.0f / 180.0f));
            if (prior_pix) {
                depth_rand = hpm_rng_uniform(&rng) * 6 * ds + (depth_prior - 3 * ds);
                n_rand = perturbed_normal(rc, px, py, prior_plane, &rng, as);
            } else {
                depth_rand = hpm_rng_uniform(&rng) * (P.depth_max - P.depth_min) + P.depth_min;
                n_rand = random_normal(rc, px, py, &rng, depth_now);
            }
            const float dmin_p = (1 - perturbation) * depth_now;
            const float dmax_p = (1 + perturbation) * depth_now;
            const float depth_pert = hpm_rng_uniform(&rng) * (dmax_p - dmin_p) + dmin_p;
            const float4 plane_now = plane_center;
            const float4 n_pert =
                perturbed_normal(rc, px, py, plane_now, &rng, (float)((double)perturbation * kPi));
            const float dnow = depth_now;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const float di = (i == 0 || i == 2) ? depth_rand : (i == 4 ? depth_pert : dnow);
                float4 t = (i == 0 || i == 4) ? plane_now : (i == 3 ? n_pert : n_rand);
                t.w = dist_to_origin(rc, px, py, di, t);
                cand[i] = t;
                cand_d[i] = di;
            }
        }
    }

    /* the refinement candidates on the selected views (HPM.cu:1191-1200),
     * views in ascending order, so the weighted sums accumulate in view
     * order. The (lane, candidate) evaluations of the lanes that selected a
     * view are packed into full trips on that view: lane j scores task
     * t0 + j, candidate t / n of the (t mod n)-th selecting lane, whose
     * plane, pixel and reference sums come by lane permutes and whose
     * window weights are in LDS; the owners take their terms back. */
    float rsum[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    {
        const int wv = (int)threadIdx.y, ln = (int)threadIdx.x; /* blockDim = (64, 4) */
        const int pxy = px | (py << 16);
#pragma unroll 1
        for (int vi = 0; vi < nsrc; ++vi) {
            const bool mine = ((sel_views >> vi) & 1u) != 0u;
            const uint64_t m = __ballot(mine);
            if (m == 0ull) {
                continue;
            }
            const int n = __popcll(m);
            const int rank = lane_rank(m);
            if (mine) {
                s_own[wv][rank] = ln;
            }
            __builtin_amdgcn_wave_barrier();
            const float wview = draws_of(vw, vi);
            const int total = 5 * n;
            /* t / n for t < 5 n <= 320: (t + 1/2) / n lies at least 1 / (2 n)
             * >= 1/128 inside an integer interval, far beyond the float
             * rounding of the product, so truncation gives the quotient */
            const float rn = 1.0f / (float)n;
            const float hrn = 0.5f * rn;
#pragma unroll 1
            for (int t0 = 0; t0 < total; t0 += 64) {
                const int t = t0 + ln;
                const bool act = t < total;
                const int c = act ? (int)fmaf((float)t, rn, hrn) : 0;
                const int owner = act ? s_own[wv][t - c * n] : ln;
                const int c_lo = t0 / n;
                const int c_hi = (t0 + 63) / n < 4 ? (t0 + 63) / n : 4;
                const int opxy = __shfl(pxy, owner);
                const RefPatch orp = shfl_ref_patch(rp, wv, owner);
                float4 plane = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int cc = 0; cc < 5; ++cc) {
                    if (cc >= c_lo && cc <= c_hi) {
                        const float4 pc = make_float4(__shfl(cand[cc].x, owner), __shfl(cand[cc].y, owner),
                                                      __shfl(cand[cc].z, owner), __shfl(cand[cc].w, owner));
                        if (c == cc) {
                            plane = pc;
                        }
                    }
                }
                float term = 0.0f;
                if (act) {
                    const int opx = opxy & 0xffff, opy = opxy >> 16;
                    GeomSample gs = {0.0f, 0.0f, 0.0f};
                    if (P.geom_consistency) {
                        gs = geom_fetch(pb, vi + 1, plane, opx, opy); /* in flight during the NCC */
                    }
                    double tq[3];
                    plane_q(pb, plane, tq);
                    const float cv = ncc(pb, orp, opx, opy, vi + 1, tq);
                    term = P.geom_consistency ? cv + gc_weight * geom_finish(pb, vi + 1, gs, opx, opy) : cv;
                }
#pragma unroll
                for (int cc = 0; cc < 5; ++cc) {
                    if (cc >= c_lo && cc <= c_hi) {
                        const int tt = cc * n + rank - t0; /* lane that scored (cc, this lane) */
                        const bool got = mine && tt >= 0 && tt < 64;
                        const float val = __shfl(term, got ? tt : ln);
                        if (got) {
                            rsum[cc] = rsum[cc] + wview * val;
                        }
                    }
                }
            }
            if (mine) {
                nev += 5u;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }

    /* the candidates decided in order (HPM.cu:1191-1218) */
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const float4 t = cand[i];
        const float tc = rsum[i] / wn;
        const float db = depth_from_plane(rc, t, px, py);
        if (prior_pix) {
            const float pr = prior_term(pk, cand_d[i] - depth_prior, prior_plane, t);
            const float rtc = hpm_expf(-tc * tc / pk.beta) * pr;
            if (db >= P.depth_min && db <= P.depth_max && rtc > restricted_cost) {
                depth_now = db;
                plane_center = t;
                plane_src = 8;
                cost_center = tc;
                restricted_cost = rtc;
            }
        } else {
            if (db >= P.depth_min && db <= P.depth_max && tc < cost_center) {
                depth_now = db;
                plane_center = t;
                plane_src = 8;
                cost_center = tc;
            }
        }
    }
    if (valid) {
        if (P.hierarchy && !(cost_center < pb.pre_cost[center] - 0.1f)) {
            /* ACMM hierarchy (ACMM.cu:1619-1624): the hypothesis stays, the
             * cost is the re-scored one, the views those of the adoption */
            pb.cost[center] = cost_now;
            pb.sel[center] = sel_center;
        } else {
            st4(pb.hyp, center, plane_center);
            pb.cost[center] = cost_center;
            pb.sel[center] = sel_center;
            if (vc != nullptr && plane_src >= 8) {
                vc[0] = __uint_as_float(0xffffffffu); /* a candidate's costs are not kept */
            } else if (vc != nullptr && plane_src >= 0) {
                /* the adopted direction's costs, eight views at a time */
#pragma unroll 1
                for (int v0 = 0; v0 < nsrc; v0 += 8) {
                    float c[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        c[k] = ca[min(v0 + k, nsrc - 1)][plane_src];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        if (v0 + k < nsrc) {
                            vc[(size_t)(v0 + k) * np] = c[k];
                        }
                    }
                }
            }
        }
    }
    return nev;
}

/* CheckerboardFilter (HPM.cu:1766-1880) */
__device__ void filter_pixel(const hpm_dev_problem& pb, int px, int py)
{
    const int W = pb.W, H = pb.H;
    const int center = py * W + px;
    if (pb.cost[center] < 0.001f) {
        return;
    }
    const float* hyp = pb.hyp;
    const int left = center - 1, leftleft = center - 3;
    const int up = center - W, upup = center - 3 * W;
    const int down = center + W, downdown = center + 3 * W;
    const int right = center + 1, rightright = center + 3;
    if (px > 4 && px < W - 5 && py > 4 && py < H - 5) {
        /* away from the border every tap exists: 21 values, and when they
         * are all positive finite depths (no NaN, no signed zeros) their
         * median is selected by a min/max sorting network in registers,
         * the same value as the insertion sort's middle element */
        const int idx[21] = {center, up, upup, upup - W * 2, down, downdown, downdown + W * 2, left, leftleft,
                             leftleft - 2, right, rightright, rightright + 2, up + 2, down + 2, up - 2, down - 2,
                             left - W * 2, right - W * 2, left + W * 2, right + W * 2};
        float v[21];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 21; ++k) {
            v[k] = hyp[4 * idx[k] + 3];
            ok = ok && v[k] > 0.0f && v[k] < INFINITY;
        }
        if (ok) {
#pragma unroll
            for (int i = 1; i < 21; ++i) {
#pragma unroll
                for (int j = i; j > 0; --j) {
                    const float lo = fminf(v[j - 1], v[j]);
                    const float hi = fmaxf(v[j - 1], v[j]);
                    v[j - 1] = lo;
                    v[j] = hi;
                }
            }
            pb.hyp[4 * center + 3] = v[10];
            return;
        }
    }
    float f[21];
    int n = 0;
    f[n++] = hyp[4 * center + 3];
#define HPM_TAP(cond, idx)                  \
    if (cond) {                             \
        f[n++] = hyp[4 * (idx) + 3];        \
    }
    HPM_TAP(py > 0, up)
    HPM_TAP(py > 2, upup)
    HPM_TAP(py > 4, upup - W * 2)
    HPM_TAP(py < H - 1, down)
    HPM_TAP(py < H - 3, downdown)
    HPM_TAP(py < H - 5, downdown + W * 2)
    HPM_TAP(px > 0, left)
    HPM_TAP(px > 2, leftleft)
    HPM_TAP(px > 4, leftleft - 2)
    HPM_TAP(px < W - 1, right)
    HPM_TAP(px < W - 3, rightright)
    HPM_TAP(px < W - 5, rightright + 2)
    HPM_TAP(py > 0 && px < W - 2, up + 2)
    HPM_TAP(py < H - 1 && px < W - 2, down + 2)
    HPM_TAP(py > 0 && px > 1, up - 2)
    HPM_TAP(py < H - 1 && px > 1, down - 2)
    HPM_TAP(px > 0 && py > 2, left - W * 2)
    HPM_TAP(px < W - 1 && py > 2, right - W * 2)
    HPM_TAP(px > 0 && py < H - 2, left + W * 2)
    HPM_TAP(px < W - 1 && py < H - 2, right + W * 2)
#undef HPM_TAP
    for (int i = 1; i < n; i++) {
        const float tmp = f[i];
        int j;
        for (j = i; j >= 1 && tmp < f[j - 1]; j--) {
            f[j] = f[j - 1];
        }
        f[j] = tmp;
    }
    const int m = n / 2;
    pb.hyp[4 * center + 3] = (n % 2 == 0) ? (f[m - 1] + f[m]) / 2 : f[m];
}

/* pixel of this thread in a colour launch: rows 2*yy + parity, black when
 * x+y is even (HPM.cu:1727-1749) */
/* Workgroups are dispatched round-robin over the 8 XCDs, each with its own
 * L2. Renumbering them so that XCD k runs the k-th contiguous band of the
 * grid keeps the source-image rows a band samples in one L2. The mapping
 * is a bijection of the grid, whatever the dispatch order. */
constexpr int kXcds = 8;

__device__ __forceinline__ void xcd_block(int* bx, int* by)
{
    const int gx = (int)gridDim.x;
    const int n = gx * (int)gridDim.y;
    const int id = (int)blockIdx.x + (int)blockIdx.y * gx;
    const int xcd = id % kXcds, i = id / kXcds;
    const int base = n / kXcds, rem = n % kXcds;
    const int l = xcd * base + min(xcd, rem) + i;
    *bx = l % gx;
    *by = l / gx;
}

/* pixel of lane (threadIdx) in tile (bx, by) of the colour grid */
__device__ __forceinline__ bool colour_pixel_at(const hpm_dev_problem& pb, int colour, int bx, int by, int* px,
                                                int* py)
{
    /* pixel (x, y) has colour (x + y) & 1; the launch of HPM.cu:1924-1931
     * visits rows y = 2 yy + parity with yy < yy_limit. Here a wave takes
     * every second pixel of HPM_WAVE_ROWS consecutive rows, so its lanes
     * sample neighbouring source columns and the windows of its rows
     * overlap. */
    const int lane = (int)threadIdx.x; /* blockDim = (64, 4) */
    const int y = (by * (int)blockDim.y + (int)threadIdx.y) * HPM_WAVE_ROWS + lane / HPM_COLOUR_TILE_X;
    const int x = 2 * (bx * HPM_COLOUR_TILE_X + lane % HPM_COLOUR_TILE_X) + ((y + colour) & 1);
    *px = x;
    *py = y;
    return (y >> 1) < pb.yy_limit && x < pb.W && y < pb.H;
}

__device__ __forceinline__ bool colour_pixel(const hpm_dev_problem& pb, int colour, int* px, int* py)
{
    int bx, by;
    xcd_block(&bx, &by);
    return colour_pixel_at(pb, colour, bx, by, px, py);
}

} // namespace

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_init(hpm_dev_problem pb)
{
    int bx, by;
    xcd_block(&bx, &by);
    const int px = bx * blockDim.x + threadIdx.x;
    const int py = by * blockDim.y + threadIdx.y;
    uint32_t nev = 0;
    if (px < pb.W && py < pb.H) {
        nev = init_pixel(pb, px, py);
    }
    count_evals(pb, nev);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_propagate(hpm_dev_problem pb, int iter, int colour)
{
    int px, py;
    const bool valid = colour_pixel(pb, colour, &px, &py);
    const uint32_t nev = propagate_pixel(pb, px, py, iter, colour, valid);
    count_evals(pb, nev);
}

/* GetDepthandNormal (HPM.cu:1751-1764) */
__device__ __forceinline__ void finalize_pixel(const hpm_dev_problem& pb, int px, int py)
{
    const hpm_camera rc = camera(pb, 0);
    const int center = py * pb.W + px;
    float4 h = ld4(pb.hyp, center);
    h.w = depth_from_plane(rc, h, px, py);
    float4 o;
    o.x = rc.R[0] * h.x + rc.R[3] * h.y + rc.R[6] * h.z;
    o.y = rc.R[1] * h.x + rc.R[4] * h.y + rc.R[7] * h.z;
    o.z = rc.R[2] * h.x + rc.R[5] * h.y + rc.R[8] * h.z;
    o.w = h.w;
    st4(pb.hyp, center, o);
}

__global__ __launch_bounds__(256) void hpm_k_finalize(hpm_dev_problem pb)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px < pb.W && py < pb.H) {
        finalize_pixel(pb, px, py);
    }
}

__global__ __launch_bounds__(256) void hpm_k_filter(hpm_dev_problem pb, int colour)
{
    int px, py;
    if (colour_pixel(pb, colour, &px, &py)) {
        filter_pixel(pb, px, py);
    }
}

/* ----------------------------------------------------------- learned prior
 * Network input and the judge of the learned depth prior (no reference
 * counterpart; the pipeline uses it in place of the planar prior passes
 * when a weight file is given, falling back to the planar prior). */

/* x[p] (32 BF16 channels): grey / 255, (depth - dmin) / (dmax - dmin),
 * camera-frame normal R n_world, cost / 2, zeros; non-finite values -> 0.
 * hyp holds (world normal, depth) of the finished plain pass. */
__global__ __launch_bounds__(256) void hpm_k_lprior_features(hpm_dev_problem pb, __bf16* x)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= pb.W || py >= pb.H) {
        return;
    }
    const hpm_camera rc = camera(pb, 0);
    const int center = py * pb.W + px;
    const float4 h = ld4(pb.hyp, center);
    const float dmin = pb.params.depth_min, dmax = pb.params.depth_max;
    float f[6];
    f[0] = pb.images[cview(pb, 0)->img_off + center] / 255.0f;
    f[1] = (h.w - dmin) / (dmax - dmin);
    f[2] = rc.R[0] * h.x + rc.R[1] * h.y + rc.R[2] * h.z;
    f[3] = rc.R[3] * h.x + rc.R[4] * h.y + rc.R[5] * h.z;
    f[4] = rc.R[6] * h.x + rc.R[7] * h.y + rc.R[8] * h.z;
    f[5] = pb.cost[center] / 2.0f;
    __bf16* o = x + (size_t)center * 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const float v = c < 6 ? f[c] : 0.0f;
        o[c] = (__bf16)(isfinite(v) ? v : 0.0f);
    }
}

/* Prediction and judge per pixel: depth = dmin + clamp(x1 + h0, 0, 1)
 * (dmax - dmin), normal = normalise(x[2..4] + h[1..3]) turned toward the
 * camera; the proposal is kept as prior (mask 1, camera-frame plane) when
 * its multi-view cost (mean of the top-k NCC, as RandomInitialization
 * scores, HPM.cu:399-436) is no worse than that of the plain pass's
 * hypothesis at the pixel. accepted[0] counts kept pixels. */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_lprior_judge(
    hpm_dev_problem pb, const __bf16* __restrict__ x, const float* __restrict__ head, float* planes, uint32_t* mask,
    unsigned long long* accepted, int use_conf)
{
    int bx, by;
    xcd_block(&bx, &by);
    const int px = bx * blockDim.x + threadIdx.x;
    const int py = by * blockDim.y + threadIdx.y;
    uint32_t keep = 0u;
    if (px < pb.W && py < pb.H) {
        const hpm_camera rc = camera(pb, 0);
        const int center = py * pb.W + px;
        const float dmin = pb.params.depth_min, dmax = pb.params.depth_max;
        const __bf16* xi = x + (size_t)center * 32;
        const float* hi = head + (size_t)center * 32;
        const float dn = fminf(fmaxf((float)xi[1] + hi[0], 0.0f), 1.0f);
        const float depth = dmin + dn * (dmax - dmin);
        float4 n = make_float4((float)xi[2] + hi[1], (float)xi[3] + hi[2], (float)xi[4] + hi[3], 0.0f);
        const float len2 = n.x * n.x + n.y * n.y + n.z * n.z;
        float4 prop = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        /* head channel 4, when trained, is the logit of the proposal being
         * right: below 0 (probability 0.5) the proposal is not judged */
        bool ok = len2 > 1e-12f && isfinite(len2) && (!use_conf || hi[4] > 0.0f);
        if (ok) {
            normalize3(&n);
            const float4 vd = view_direction(rc, px, py, depth);
            if (dot3(n, vd) > 0.0f) {
                n.x = -n.x;
                n.y = -n.y;
                n.z = -n.z;
            }
            prop = n;
            prop.w = dist_to_origin(rc, px, py, depth, prop);
            /* current hypothesis: (world normal, depth) -> camera plane */
            const float4 t = ld4(pb.hyp, center);
            float4 cur;
            cur.x = rc.R[0] * t.x + rc.R[1] * t.y + rc.R[2] * t.z;
            cur.y = rc.R[3] * t.x + rc.R[4] * t.y + rc.R[5] * t.z;
            cur.z = rc.R[6] * t.x + rc.R[7] * t.y + rc.R[8] * t.z;
            cur.w = dist_to_origin(rc, px, py, t.w, cur);
            RefPatch rp;
            make_ref_patch(pb.images + cview(pb, 0)->img_off, pb.W, pb.H, px, py, pb.params, &rp);
            uint32_t sel, nev = 0;
            const float c_prop = initial_cost(pb, rp, px, py, prop, &sel, &nev);
            const float c_cur = isfinite(t.w) ? initial_cost(pb, rp, px, py, cur, &sel, &nev) : 2.0f;
            ok = c_prop <= c_cur && depth >= dmin && depth <= dmax;
        }
        st4(planes, center, ok ? prop : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        mask[center] = ok ? 1u : 0u;
        keep = ok ? 1u : 0u;
    }
    const unsigned long long s = wave_sum((unsigned long long)keep);
    if ((threadIdx.x + blockDim.x * threadIdx.y) % 64 == 0 && s != 0ull) {
        atomicAdd(accepted, s);
    }
}

/* ---------------------------------------------------------- batched form
 * Several independent problems in one launch per stage: a grid of resident
 * blocks walks a table of tiles (problem, tile x, tile y). Two schedules,
 * chosen by the autotuner:
 *  - static: logical block l (XCD-renumbered) takes the contiguous tile
 *    range [l n / G, (l + 1) n / G), so each XCD works through a
 *    contiguous stretch of problems and rows;
 *  - work queue: the tiles are cut into 8 contiguous bands, one per XCD;
 *    a block pulls work items of kGran consecutive tiles from its XCD's
 *    band with an atomic counter and, once that band is empty, from the
 *    next XCDs' bands in turn, so uneven tiles balance while each XCD's L2
 *    keeps its band. Every block leaves once every band is empty.
 * Per pixel the work is the unbatched kernels' own. */

__device__ __forceinline__ hpm_tile load_tile(const hpm_tile* tiles, int t)
{
    typedef const __attribute__((address_space(4))) hpm_tile* ctile;
    const ctile p = (ctile)tiles + t;
    hpm_tile r;
    r.problem = p->problem;
    r.bx = p->bx;
    r.by = p->by;
    r.pad_ = 0;
    return r;
}

__device__ __forceinline__ void tile_range(int ntiles, int* first, int* last)
{
    const int g = (int)gridDim.x;
    const int xcd = (int)blockIdx.x % kXcds, i = (int)blockIdx.x / kXcds;
    const int base = g / kXcds, rem = g % kXcds;
    const int l = xcd * base + min(xcd, rem) + i;
    *first = (int)((int64_t)l * ntiles / g);
    *last = (int)((int64_t)(l + 1) * ntiles / g);
}

/* Calls body(t) for every tile t this block takes under the schedule:
 * kGran == 0 the static ranges, else work items of kGran tiles from the
 * per-XCD queues (queue: kXcds zeroed counters of this launch). The loop
 * condition is uniform across the block (the item goes through LDS). */
template <int kGran, typename F>
__device__ __forceinline__ void for_tiles(int ntiles, unsigned int* queue, F&& body)
{
    if (kGran == 0) {
        int first, last;
        tile_range(ntiles, &first, &last);
        for (int t = first; t < last; ++t) {
            body(t);
        }
        return;
    }
    __shared__ int s_item;
    const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
    const int home = (int)blockIdx.x % kXcds;
    for (int k = 0; k < kXcds; ++k) {
        const int band = (home + k) % kXcds;
        const int b0 = (int)((int64_t)band * ntiles / kXcds);
        const int b1 = (int)((int64_t)(band + 1) * ntiles / kXcds);
        const int items = (b1 - b0 + kGran - 1) / kGran;
        for (;;) {
            __syncthreads(); /* the previous item's s_item reads are done */
            if (leader) {
                s_item = (int)atomicAdd(&queue[band], 1u);
            }
            __syncthreads();
            const int item = s_item;
            if (item >= items) {
                break;
            }
            const int t0 = b0 + item * kGran;
            const int t1 = min(t0 + kGran, b1);
            for (int t = t0; t < t1; ++t) {
                body(t);
            }
        }
    }
}

template <int kGran>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_init_batch(
    const hpm_dev_problem* __restrict__ pbs, const hpm_tile* __restrict__ tiles, int ntiles, unsigned int* queue)
{
    for_tiles<kGran>(ntiles, queue, [&](int t) {
        const hpm_tile tl = load_tile(tiles, t);
        const hpm_dev_problem& pb = pbs[tl.problem];
        const int px = tl.bx * blockDim.x + threadIdx.x;
        const int py = tl.by * blockDim.y + threadIdx.y;
        uint32_t nev = 0;
        if (px < pb.W && py < pb.H) {
            nev = init_pixel(pb, px, py);
        }
        count_evals(pb, nev);
    });
}

template <int kGran>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_propagate_batch(
    const hpm_dev_problem* __restrict__ pbs, const hpm_tile* __restrict__ tiles, int ntiles, int iter, int colour,
    unsigned int* queue)
{
    for_tiles<kGran>(ntiles, queue, [&](int t) {
        const hpm_tile tl = load_tile(tiles, t);
        const hpm_dev_problem& pb = pbs[tl.problem];
        int px, py;
        const bool valid = colour_pixel_at(pb, colour, tl.bx, tl.by, &px, &py);
        const uint32_t nev = propagate_pixel(pb, px, py, iter, colour, valid);
        count_evals(pb, nev);
    });
}

template <int kGran>
__global__ __launch_bounds__(256) void hpm_k_finalize_batch(const hpm_dev_problem* __restrict__ pbs,
                                                            const hpm_tile* __restrict__ tiles, int ntiles,
                                                            unsigned int* queue)
{
    for_tiles<kGran>(ntiles, queue, [&](int t) {
        const hpm_tile tl = load_tile(tiles, t);
        const hpm_dev_problem& pb = pbs[tl.problem];
        const int px = tl.bx * blockDim.x + threadIdx.x;
        const int py = tl.by * blockDim.y + threadIdx.y;
        if (px < pb.W && py < pb.H) {
            finalize_pixel(pb, px, py);
        }
    });
}

template <int kGran>
__global__ __launch_bounds__(256) void hpm_k_filter_batch(const hpm_dev_problem* __restrict__ pbs,
                                                          const hpm_tile* __restrict__ tiles, int ntiles, int colour,
                                                          unsigned int* queue)
{
    for_tiles<kGran>(ntiles, queue, [&](int t) {
        const hpm_tile tl = load_tile(tiles, t);
        const hpm_dev_problem& pb = pbs[tl.problem];
        int px, py;
        if (colour_pixel_at(pb, colour, tl.bx, tl.by, &px, &py)) {
            filter_pixel(pb, px, py);
        }
    });
}

/* The whole batched RunPatchMatch (init, 3 x red/black propagation,
 * finalize, red/black filter) in one persistent cooperative launch: every
 * block stays resident, pulls work items of kGran tiles from the per-XCD
 * queues of each stage (queue: 8 zeroed counters per stage), and the grid
 * synchronises between stages, whose reads depend on the other colour's or
 * the previous stage's writes anywhere in the problem. The grid barrier
 * makes those writes visible across XCDs. */
template <int kGran>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_batch_persistent(
    const hpm_dev_problem* __restrict__ pbs, int n, const hpm_tile* __restrict__ tiles, int n_full, int n_colour,
    unsigned int* queue)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const hpm_dev_problem* p_init = pbs;
    const hpm_dev_problem* p_prop = pbs + n;
    const hpm_tile* tf = tiles;
    const hpm_tile* tc = tiles + n_full;
    for_tiles<kGran>(n_full, queue, [&](int t) {
        const hpm_tile tl = load_tile(tf, t);
        const hpm_dev_problem& pb = p_init[tl.problem];
        const int px = tl.bx * blockDim.x + threadIdx.x;
        const int py = tl.by * blockDim.y + threadIdx.y;
        uint32_t nev = 0;
        if (px < pb.W && py < pb.H) {
            nev = init_pixel(pb, px, py);
        }
        count_evals(pb, nev);
    });
    grid.sync();
#pragma unroll 1
    for (int k = 0; k < 6; ++k) { /* 3 iterations (HPM.cu:1933) x black, red */
        const int iter = k / 2, colour = k % 2;
        for_tiles<kGran>(n_colour, queue + 8 * (1 + k), [&](int t) {
            const hpm_tile tl = load_tile(tc, t);
            const hpm_dev_problem& pb = p_prop[tl.problem];
            int px, py;
            const bool valid = colour_pixel_at(pb, colour, tl.bx, tl.by, &px, &py);
            const uint32_t nev = propagate_pixel(pb, px, py, iter, colour, valid);
            count_evals(pb, nev);
        });
        grid.sync();
    }
    for_tiles<kGran>(n_full, queue + 8 * 7, [&](int t) {
        const hpm_tile tl = load_tile(tf, t);
        const hpm_dev_problem& pb = p_prop[tl.problem];
        const int px = tl.bx * blockDim.x + threadIdx.x;
        const int py = tl.by * blockDim.y + threadIdx.y;
        if (px < pb.W && py < pb.H) {
            finalize_pixel(pb, px, py);
        }
    });
#pragma unroll 1
    for (int colour = 0; colour < 2; ++colour) {
        grid.sync();
        for_tiles<kGran>(n_colour, queue + 8 * (8 + colour), [&](int t) {
            const hpm_tile tl = load_tile(tc, t);
            const hpm_dev_problem& pb = p_prop[tl.problem];
            int px, py;
            if (colour_pixel_at(pb, colour, tl.bx, tl.by, &px, &py)) {
                filter_pixel(pb, px, py);
            }
        });
    }
}

/* zeroes n counters (a kernel, so the schedule can be captured in a graph) */
__global__ __launch_bounds__(64) void hpm_k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0ull;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HPM_NCC_WAVES))) void hpm_k_cost_vectors(hpm_dev_problem pb, const float* planes,
                                                          const int32_t* pix, int npix, float* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) {
        return;
    }
    const int px = pix[2 * i], py = pix[2 * i + 1];
    const float4 pl = ld4(planes, i);
    RefPatch rp;
    make_ref_patch(pb.images + cview(pb, 0)->img_off, pb.W, pb.H, px, py, pb.params, &rp);
    double q[3];
    plane_q(pb, pl, q);
    for (int v = 1; v <= pb.nsrc; ++v) {
        out[(size_t)i * pb.nsrc + v - 1] = ncc(pb, rp, px, py, v, q);
    }
}

__global__ __launch_bounds__(256) void hpm_k_jbu(const float* guide, int w, int h, const float* src_nd, int sw,
                                                 int sh, int scale_i, float* out_depth, float* out_normal)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= w || py >= h) {
        return;
    }
    float d;
    float4 n;
    jbu_pixel(guide, w, h, src_nd, sw, sh, scale_i, px, py, &d, &n);
    const int center = py * w + px;
    out_depth[center] = d;
    st4(out_normal, center, n);
}

/* JBU followed by the prior reload of main.cpp:179-204 (scale 0) and
 * main.cpp:330-353 (scale 1): TransformNormal2RefCam (HPM.cpp:882-890) and
 * GetDistance2Origin_factor (HPM.cpp:892-899, full-resolution K) */
__global__ __launch_bounds__(256) void hpm_k_jbu_prior(const float* guide, int w, int h, const float* src_nd, int sw,
                                                       int sh, int scale_i, hpm_camera ref, float depth_min,
                                                       float depth_max, float* planes, uint32_t* mask)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= w || py >= h) {
        return;
    }
    float d;
    float4 n;
    jbu_pixel(guide, w, h, src_nd, sw, sh, scale_i, px, py, &d, &n);
    const int center = py * w + px;
    uint32_t m = 0u;
    float4 pl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (d <= depth_max && d >= depth_min) {
        m = 1u;
        pl.x = ref.R[0] * n.x + ref.R[1] * n.y + ref.R[2] * n.z;
        pl.y = ref.R[3] * n.x + ref.R[4] * n.y + ref.R[5] * n.z;
        pl.z = ref.R[6] * n.x + ref.R[7] * n.y + ref.R[8] * n.z;
        pl.w = dist_to_origin(ref, px, py, d, pl);
    }
    mask[center] = m;
    st4(planes, center, pl);
}

__global__ __launch_bounds__(256) void hpm_k_store_state(const float* hyp, const float* cost, int64_t n,
                                                         float* out_hyp, float* out_cost, float* out_depth)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float4 v = reinterpret_cast<const float4*>(hyp)[i];
    reinterpret_cast<float4*>(out_hyp)[i] = v;
    out_cost[i] = cost[i];
    out_depth[i] = v.w;
}

__global__ __launch_bounds__(256) void hpm_k_selftest(int kind, const void* in, void* out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    if (kind == 4) {
        const uint32_t* b = static_cast<const uint32_t*>(in) + 6 * (size_t)i;
        const uint32_t ctr[4] = {b[0], b[1], b[2], b[3]};
        const uint32_t key[2] = {b[4], b[5]};
        uint32_t r[4];
        hpm_philox4x32_10(ctr, key, r);
        uint32_t* o = static_cast<uint32_t*>(out) + 4 * (size_t)i;
        o[0] = r[0];
        o[1] = r[1];
        o[2] = r[2];
        o[3] = r[3];
        return;
    }
    if (kind == 8) {
        const float* b = static_cast<const float*>(in) + 2 * (size_t)i;
        static_cast<float*>(out)[i] = div_rn(b[0], b[1]);
        return;
    }
    if (kind == 9) {
        static_cast<double*>(out)[i] = rcp_rn(static_cast<const double*>(in)[i]);
        return;
    }
    const float x = static_cast<const float*>(in)[i];
    float y;
    switch (kind) {
    case 7: y = sqrt_rn(x); break;
    case 0: y = hpm_expf(x); break;
    case 1: y = hpm_sinf(x); break;
    case 2: y = hpm_cosf(x); break;
    case 5: y = fract1(x); break;
    case 6: y = __int_as_float(floor_int(x)); break;
    default: y = hpm_acosf(x); break;
    }
    static_cast<float*>(out)[i] = y;
}

/* ---- launchers ---- */

static dim3 full_grid(int W, int H, dim3 blk)
{
    return dim3((W + blk.x - 1) / blk.x, (H + blk.y - 1) / blk.y, 1);
}

static dim3 colour_grid(const hpm_dev_problem* pb)
{
    const int half_w = (pb->W + 1) / 2;
    const int rows = pb->H < 2 * pb->yy_limit ? pb->H : 2 * pb->yy_limit;
    return dim3((half_w + HPM_COLOUR_TILE_X - 1) / HPM_COLOUR_TILE_X,
                (rows + HPM_COLOUR_TILE_Y - 1) / HPM_COLOUR_TILE_Y, 1);
}

extern "C" int hpm_launch_init(const hpm_dev_problem* pb, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_init, full_grid(pb->W, pb->H, blk), blk, 0, (hipStream_t)stream, *pb);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_propagate(const hpm_dev_problem* pb, int iter, int colour, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_propagate, colour_grid(pb), blk, 0, (hipStream_t)stream, *pb, iter, colour);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_finalize(const hpm_dev_problem* pb, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_finalize, full_grid(pb->W, pb->H, blk), blk, 0, (hipStream_t)stream, *pb);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_filter(const hpm_dev_problem* pb, int colour, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_filter, colour_grid(pb), blk, 0, (hipStream_t)stream, *pb, colour);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_cost_vectors(const hpm_dev_problem* pb, const float* planes, const int32_t* pix,
                                       int npix, float* out, void* stream)
{
    if (npix <= 0) {
        return 0;
    }
    const dim3 blk(256, 1, 1);
    const dim3 grd((npix + 255) / 256, 1, 1);
    hipLaunchKernelGGL(hpm_k_cost_vectors, grd, blk, 0, (hipStream_t)stream, *pb, planes, pix, npix, out);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_jbu(const float* guide, int w, int h, const float* src_nd, int sw, int sh, int scale_i,
                              float* out_depth, float* out_normal, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_jbu, full_grid(w, h, blk), blk, 0, (hipStream_t)stream, guide, w, h, src_nd, sw, sh,
                       scale_i, out_depth, out_normal);
    return (int)hipGetLastError();
}

namespace {

template <int kGran>
void launch_batch_stage(int stage, dim3 g, dim3 blk, hipStream_t s, const hpm_dev_problem* pbs,
                        const hpm_tile* tiles, int ntiles, int iter, int colour, unsigned int* queue)
{
    switch (stage) {
    case HPM_STAGE_INIT:
        hipLaunchKernelGGL(hpm_k_init_batch<kGran>, g, blk, 0, s, pbs, tiles, ntiles, queue);
        break;
    case HPM_STAGE_PROPAGATE:
        hipLaunchKernelGGL(hpm_k_propagate_batch<kGran>, g, blk, 0, s, pbs, tiles, ntiles, iter, colour, queue);
        break;
    case HPM_STAGE_FINALIZE:
        hipLaunchKernelGGL(hpm_k_finalize_batch<kGran>, g, blk, 0, s, pbs, tiles, ntiles, queue);
        break;
    default:
        hipLaunchKernelGGL(hpm_k_filter_batch<kGran>, g, blk, 0, s, pbs, tiles, ntiles, colour, queue);
        break;
    }
}

} // namespace

extern "C" int hpm_launch_batch(int stage, const hpm_dev_problem* pbs, const hpm_tile* tiles, int ntiles, int grid,
                                int iter, int colour, int gran, unsigned int* queue, void* stream)
{
    if (ntiles <= 0) {
        return 0;
    }
    if (stage < HPM_STAGE_INIT || stage > HPM_STAGE_FILTER || (gran != 0 && !queue)) {
        return (int)hipErrorInvalidValue;
    }
    const dim3 blk(64, 4, 1);
    const dim3 g(grid > 0 ? grid : 1, 1, 1);
    const hipStream_t s = (hipStream_t)stream;
    switch (gran) {
    case 0: launch_batch_stage<0>(stage, g, blk, s, pbs, tiles, ntiles, iter, colour, queue); break;
    case 1: launch_batch_stage<1>(stage, g, blk, s, pbs, tiles, ntiles, iter, colour, queue); break;
    case 2: launch_batch_stage<2>(stage, g, blk, s, pbs, tiles, ntiles, iter, colour, queue); break;
    case 4: launch_batch_stage<4>(stage, g, blk, s, pbs, tiles, ntiles, iter, colour, queue); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

namespace {

template <int kGran>
const void* persistent_kernel()
{
    return (const void*)hpm_k_batch_persistent<kGran>;
}

} // namespace

extern "C" int hpm_batch_persistent_grid(int device, int gran, int* grid)
{
    const void* k = gran == 1 ? persistent_kernel<1>() : gran == 2 ? persistent_kernel<2>() : nullptr;
    if (!k || !grid) {
        return (int)hipErrorInvalidValue;
    }
    int per_cu = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    *grid = per_cu * cus;
    return (int)e;
}

extern "C" int hpm_launch_batch_persistent(const hpm_dev_problem* pbs, int n, const hpm_tile* tiles, int n_full,
                                           int n_colour, int gran, unsigned int* queue, int grid, void* stream)
{
    const void* k = gran == 1 ? persistent_kernel<1>() : gran == 2 ? persistent_kernel<2>() : nullptr;
    if (!k || !queue || grid < 1) {
        return (int)hipErrorInvalidValue;
    }
    void* args[] = {(void*)&pbs, (void*)&n, (void*)&tiles, (void*)&n_full, (void*)&n_colour, (void*)&queue};
    return (int)hipLaunchCooperativeKernel(k, dim3(grid, 1, 1), dim3(64, 4, 1), args, 0, (hipStream_t)stream);
}

extern "C" int hpm_launch_lprior_features(const hpm_dev_problem* pb, void* x, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_lprior_features, full_grid(pb->W, pb->H, blk), blk, 0, (hipStream_t)stream, *pb,
                       (__bf16*)x);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_lprior_judge(const hpm_dev_problem* pb, const void* x, const float* head, float* planes,
                                       uint32_t* mask, unsigned long long* accepted, int use_conf, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_lprior_judge, full_grid(pb->W, pb->H, blk), blk, 0, (hipStream_t)stream, *pb,
                       (const __bf16*)x, head, planes, mask, accepted, use_conf);
    return (int)hipGetLastError();
}

/* ---- fusion (main.cpp:502-633): per (pixel, source) consistency ---- */

namespace {

struct P3 {
    float x, y, z;
};

/* Get3DPointonWorld (HPM.cpp:209-233) */
__device__ __forceinline__ P3 fusion_world_point(int x, int y, float depth, const hpm_camera& c)
{
    const float px = depth * ((float)x - c.K[2]) / c.K[0];
    const float py = depth * ((float)y - c.K[5]) / c.K[4];
    const float pz = depth;
    const float tx = c.R[0] * px + c.R[3] * py + c.R[6] * pz;
    const float ty = c.R[1] * px + c.R[4] * py + c.R[7] * pz;
    const float tz = c.R[2] * px + c.R[5] * py + c.R[8] * pz;
    const float cx = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
    const float cy = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
    const float cz = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
    return P3{tx + cx, ty + cy, tz + cz};
}

/* ProjectonCamera (HPM.cpp:246-256) */
__device__ __forceinline__ void fusion_project(const P3& X, const hpm_camera& c, float* u, float* v, float* depth)
{
    const float tx = c.R[0] * X.x + c.R[1] * X.y + c.R[2] * X.z + c.t[0];
    const float ty = c.R[3] * X.x + c.R[4] * X.y + c.R[5] * X.z + c.t[1];
    const float tz = c.R[6] * X.x + c.R[7] * X.y + c.R[8] * X.z + c.t[2];
    *depth = c.K[6] * tx + c.K[7] * ty + c.K[8] * tz;
    *u = (c.K[0] * tx + c.K[1] * ty + c.K[2] * tz) / *depth;
    *v = (c.K[3] * tx + c.K[4] * ty + c.K[5] * tz) / *depth;
}

} // namespace

__global__ __launch_bounds__(256) void hpm_k_fusion_geometry(hpm_fusion_args a, int width, int height)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int r = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (c >= width || r >= height) {
        return;
    }
    const size_t p = (size_t)r * width + c;
    const hpm_camera rc = a.cams[a.ref];
    const float ref_depth = a.depth[a.off[a.ref] + p];
    const float4 rn = reinterpret_cast<const float4*>(a.hyp)[a.off[a.ref] + p];
    uint32_t any = 0u;
    for (int j = 0; j < a.nn; ++j) {
        int32_t pix = -1;
        float w = 0.0f;
        if (ref_depth > 0.0f) {
            const P3 X = fusion_world_point(c, r, ref_depth, rc);
            const int s = a.src[j];
            const hpm_camera sc = a.cams[s];
            float u, v, proj_depth;
            fusion_project(X, sc, &u, &v, &proj_depth);
            const int sr = (int)(v + 0.5f), scol = (int)(u + 0.5f); /* truncation, as int(x + 0.5f) */
            if (scol >= 0 && scol < sc.width && sr >= 0 && sr < sc.height) {
                const size_t q = (size_t)sr * sc.width + scol;
                const float src_depth = a.depth[a.off[s] + q];
                if (src_depth > 0.0f) {
                    const float4 sn = reinterpret_cast<const float4*>(a.hyp)[a.off[s] + q];
                    const P3 Y = fusion_world_point(scol, sr, src_depth, sc);
                    float tu, tv;
                    fusion_project(Y, rc, &tu, &tv, &proj_depth);
                    const double dx = (double)((float)c - tu), dy = (double)((float)r - tv);
                    const float reproj = (float)sqrt(dx * dx + dy * dy);
                    const float rdd = fabsf(proj_depth - ref_depth) / ref_depth;
                    const float dot = rn.x * sn.x + rn.y * sn.y + rn.z * sn.z;
                    float angle = hpm_acosf(dot);
                    angle = angle != angle ? 0.0f : angle; /* GetAngle, HPM.cpp:258-268 */
                    if (reproj < 2.0f && rdd < 0.01f && angle < 0.174533f) {
                        pix = (int32_t)q;
                        w = hpm_expf(-(reproj + 200.0f * rdd + angle * 10.0f));
                        any |= 1u << j;
                    }
                }
            }
        }
        a.rec_pix[p * a.nn + j] = pix;
        a.rec_w[p * a.nn + j] = w;
    }
    a.any[p] = any;
}

extern "C" int hpm_launch_fusion_geometry(const hpm_fusion_args* a, int width, int height, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_fusion_geometry, full_grid(width, height, blk), blk, 0, (hipStream_t)stream, *a, width,
                       height);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_zero_u64(unsigned long long* p, int n, void* stream)
{
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(hpm_k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_selftest(int kind, const void* in, void* out, int n, void* stream)
{
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(hpm_k_selftest, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, kind, in, out, n);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_store_state(const float* hyp, const float* cost, int64_t n, float* out_hyp, float* out_cost,
                                      float* out_depth, void* stream)
{
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(hpm_k_store_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, hyp,
                       cost, n, out_hyp, out_cost, out_depth);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_jbu_prior(const float* guide, int w, int h, const float* src_nd, int sw, int sh, int scale_i,
                                    const hpm_camera* ref, float depth_min, float depth_max, float* planes,
                                    uint32_t* mask, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_jbu_prior, full_grid(w, h, blk), blk, 0, (hipStream_t)stream, guide, w, h, src_nd, sw,
                       sh, scale_i, *ref, depth_min, depth_max, planes, mask);
    return (int)hipGetLastError();
}
