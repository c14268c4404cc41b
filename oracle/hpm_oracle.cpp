/*
 * hpm_oracle.cpp — CPU (OpenMP) restatement of HPM-MVS's per-reference-view
 * PatchMatch. TEST INFRASTRUCTURE: the checker for the HIP path, never part
 * of the product.
 *
 * Every function cites the reference text it restates. Where the reference
 * depends on CUDA hardware behaviour that cannot be reproduced (fast-math
 * exp/sin/cos/rsqrt, 9-bit texture filter weights, cuRAND XORWOW seeded from
 * clock64()), this file and the HIP kernels implement one shared definition
 * instead:
 *   - exp/sin/cos/acos: hpm_fmath.h; 1/sqrt(x) as 1.0f / sqrtf(x);
 *   - bilinear sampling as (c0 + a c1) + b (c2 + a c3) of the footprint's
 *     coefficients (c0 = t00, c1 = t10 - t00, c2 = t01 - t00,
 *     c3 = t11 - t10 - t01 + t00, hpm_half.h), three fused multiply-adds
 *     with exact FP32 fractions, clamp-to-edge addressing, texel centres at
 *     integer coordinates (the reference samples at +0.5);
 *   - random numbers: Philox4x32-10 (hpm_philox.h), one counter stream per
 *     (pixel, launch);
 *   - the homography and the warped window centre in FP64 per (hypothesis,
 *     source view), the 36 tap offsets from the centre in FP32 with one
 *     reciprocal 1 / (1 + u) per tap: the linear 1 - u when the window
 *     bounds |u| by 0.002 (a tap moves by at most 4.1e-6 of its offset),
 *     else the cubic 1 - u + u^2 - u^3 when it bounds |u| by 0.01 (below
 *     1.02e-8 relative, under half an FP32 ulp), the correctly rounded
 *     quotient otherwise; bilinear fractions x - floor(x) capped at the largest float
 *     below 1 (they differ from the plain difference only where it rounds
 *     up to 1, for x just below an integer);
 *   - NCC window sums of both signals accumulated relative to the reference
 *     centre value: NCC is invariant to a shift of either signal, and the
 *     shifted sums keep FP32 accurate on low-texture patches;
 *   - source images sampled through FP16 coefficients of their FP16-rounded
 *     texels (hpm_half.h; exact for 8-bit grey images), the reference image
 *     at full FP32.
 * The file must be compiled with -ffp-contract=off and without fast-math.
 */
#include "hpm_oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <omp.h>

#include "hpm_fmath.h"
#include "hpm_fusion.h"
#include "hpm_half.h"
#include "hpm_philox.h"

namespace {

struct F4 {
    float x, y, z, w;
};

static inline F4 ld4(const float* a, int i)
{
    F4 r = {a[4 * i + 0], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]};
    return r;
}

static inline void st4(float* a, int i, F4 v)
{
    a[4 * i + 0] = v.x;
    a[4 * i + 1] = v.y;
    a[4 * i + 2] = v.z;
    a[4 * i + 3] = v.w;
}

static inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

static const double kPi = 3.14159265358979323846; /* main.h:41 */

/* Per-iteration thresholds. JudgeExtend (HPM.cu:1040-1041) evaluates
 * 0.8*exp(-iter^2*(3-k)/90) in double; view selection (HPM.cu:1568,1586)
 * evaluates 0.8*expf(iter^2/-90) and expf(thr^2/-0.32) in float. */
struct IterConsts {
    float good[3][3];   /* [iter][k] */
    float thr[3];       /* view-selection cost threshold */
    float fallback[3];  /* expf(thr^2 / -0.32) */
};

static IterConsts make_iter_consts()
{
    IterConsts c;
    for (int it = 0; it < 3; ++it) {
        for (int k = 0; k < 3; ++k) {
            const int iter_tmp = 3 - k;
            c.good[it][k] = (float)(0.8 * exp((double)(-it * it * iter_tmp) / 90.0));
        }
        const float e = hpm_expf((float)(it * it) / (-90.0f));
        c.thr[it] = (float)(0.8 * (double)e);
        c.fallback[it] = hpm_expf(c.thr[it] * c.thr[it] / (-0.32f));
    }
    return c;
}

/* Plane-induced homography (ComputeHomography, HPM.cu:227-287),
 * H = Ks (Rrel - trel n^T / w) Kr^-1 with Ks = [sK0 0 sK2; 0 sK4 sK5; 0 0 sK8]
 * and Kr = [fx 0 cx; 0 fy cy; 0 0 1] as the reference reads them, split as
 * H = A - b q^T: A = Ks Rrel Kr^-1 and b = Ks trel per source view (double,
 * from the float cameras), q = Kr^-T n / w per hypothesis. */
struct ViewRel {
    double A[9];
    double b[3];
};

struct RefK {
    double ifx, ify, cx, cy;
};

static RefK make_ref_k(const hpm_camera& rc)
{
    RefK k;
    k.ifx = 1.0 / (double)rc.K[0];
    k.ify = 1.0 / (double)rc.K[4];
    k.cx = (double)rc.K[2];
    k.cy = (double)rc.K[5];
    return k;
}

static ViewRel make_view_rel(const hpm_camera& rc, const hpm_camera& sc)
{
    double rR[9], sR[9], refC[3], srcC[3], Crel[3], Rrel[9], trel[3], M[9];
    for (int i = 0; i < 9; ++i) {
        rR[i] = (double)rc.R[i];
        sR[i] = (double)sc.R[i];
    }
    for (int a = 0; a < 3; ++a) {
        refC[a] = -(rR[a] * (double)rc.t[0] + rR[3 + a] * (double)rc.t[1] + rR[6 + a] * (double)rc.t[2]);
        srcC[a] = -(sR[a] * (double)sc.t[0] + sR[3 + a] * (double)sc.t[1] + sR[6 + a] * (double)sc.t[2]);
    }
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            Rrel[3 * r + c] = sR[3 * r] * rR[3 * c] + sR[3 * r + 1] * rR[3 * c + 1] + sR[3 * r + 2] * rR[3 * c + 2];
        }
    }
    for (int a = 0; a < 3; ++a) {
        Crel[a] = refC[a] - srcC[a];
    }
    for (int r = 0; r < 3; ++r) {
        trel[r] = sR[3 * r] * Crel[0] + sR[3 * r + 1] * Crel[1] + sR[3 * r + 2] * Crel[2];
    }
    const double sK0 = sc.K[0], sK2 = sc.K[2], sK4 = sc.K[4], sK5 = sc.K[5], sK8 = sc.K[8];
    for (int c = 0; c < 3; ++c) {
        M[c] = sK0 * Rrel[c] + sK2 * Rrel[6 + c];
        M[3 + c] = sK4 * Rrel[3 + c] + sK5 * Rrel[6 + c];
        M[6 + c] = sK8 * Rrel[6 + c];
    }
    const RefK k = make_ref_k(rc);
    ViewRel v;
    for (int r = 0; r < 3; ++r) {
        v.A[3 * r + 0] = M[3 * r + 0] * k.ifx;
        v.A[3 * r + 1] = M[3 * r + 1] * k.ify;
        v.A[3 * r + 2] = M[3 * r + 2] - M[3 * r + 0] * k.cx * k.ifx - M[3 * r + 1] * k.cy * k.ify;
    }
    v.b[0] = sK0 * trel[0] + sK2 * trel[2];
    v.b[1] = sK4 * trel[1] + sK5 * trel[2];
    v.b[2] = sK8 * trel[2];
    return v;
}

/* q = Kr^-T n / w for one hypothesis (double) */
static inline void plane_q(const RefK& k, F4 n, double q[3])
{
    const double iw = 1.0 / (double)n.w;
    const double nx = n.x, ny = n.y, nz = n.z;
    q[0] = (nx * k.ifx) * iw;
    q[1] = (ny * k.ify) * iw;
    q[2] = ((nz - (nx * k.cx) * k.ifx) - (ny * k.cy) * k.ify) * iw;
}

/* The warp of one (hypothesis, source view) around reference pixel p:
 * the warped centre s0 = H p (ComputeCorrespondingPoint, HPM.cu:289-296)
 * in double, split into an integer texel (ix0, iy0) and a float fraction,
 * and the tap offsets of s(p + d) - s0 as a float rational function of d:
 *   s(p + d) - s0 = (g d) / (1 + e . d). */
struct Warp {
    int ix0, iy0;
    float fx0, fy0;
    float ex, ey, gxx, gxy, gyx, gyy;
    bool small; /* |e . d| <= 0.01 over the window: cubic reciprocal */
    bool tiny;  /* |e . d| <= 0.002 over the window: linear reciprocal */
};

/* The reciprocal tiers of a warp, radius the largest tap offset of the
 * window: with |u| <= bu = radius (|ex| + |ey|), the linear 1 - u errs from
 * 1 / (1 + u) by u^2 / (1 + u), so a tap moves by at most bu^2 / (1 - bu)
 * of its offset (4.1e-6 at bu = 0.002). */
static inline void set_tiers(Warp* w, float radius)
{
    const float bu = (fabsf(w->ex) + fabsf(w->ey)) * radius;
    w->small = bu <= 0.01f;
    w->tiny = bu <= 0.002f;
}

/* Returns false when the warped centre falls outside the source image (the
 * reference returns the maximum cost there, HPM.cu:333-335). radius is the
 * largest tap offset of the window. */
static inline bool make_warp(const ViewRel& v, const double q[3], int px, int py, int SW, int SH, float radius,
                             Warp* w)
{
    double H[9];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            H[3 * r + c] = fma(-v.b[r], q[c], v.A[3 * r + c]);
        }
    }
    const double X = (double)px, Y = (double)py;
    const double hx = fma(H[0], X, fma(H[1], Y, H[2]));
    const double hy = fma(H[3], X, fma(H[4], Y, H[5]));
    const double hz = fma(H[6], X, fma(H[7], Y, H[8]));
    const double icz = 1.0 / hz;
    const double sx0 = hx * icz;
    const double sy0 = hy * icz;
    if (!(sx0 >= 0.0 && sx0 < (double)SW && sy0 >= 0.0 && sy0 < (double)SH)) {
        return false;
    }
    const double fx = floor(sx0), fy = floor(sy0);
    w->ix0 = (int)fx;
    w->iy0 = (int)fy;
    w->fx0 = (float)(sx0 - fx);
    w->fy0 = (float)(sy0 - fy);
    w->ex = (float)(H[6] * icz);
    w->ey = (float)(H[7] * icz);
    w->gxx = (float)(fma(-sx0, H[6], H[0]) * icz);
    w->gxy = (float)(fma(-sx0, H[7], H[1]) * icz);
    w->gyx = (float)(fma(-sy0, H[6], H[3]) * icz);
    w->gyy = (float)(fma(-sy0, H[7], H[4]) * icz);
    set_tiers(w, radius);
    return true;
}

/* 1 / (1 + u) of a tap (u = ex dx + ey dy): when the warp bounds |u| by
 * 0.002, the linear 1 - ex dx - ey dy as two fused multiply-adds (the
 * column term 1 - ex dx first); else the cubic 1 - u + u^2 - u^3 in Horner
 * form when it bounds |u| by 0.01 (u^4 / (1 + u) <= 1.02e-8), else the
 * correctly rounded quotient. */
static inline float tap_inv(float u, const Warp& w, float dx, float dy)
{
    if (w.tiny) {
        return fmaf(-w.ey, dy, fmaf(-w.ex, dx, 1.0f));
    }
    return w.small ? fmaf(u, fmaf(u, 1.0f - u, -1.0f), 1.0f) : 1.0f / (1.0f + u);
}

/* The bilinear fraction of x: x - floor(x), capped at the largest float
 * below 1 (the difference rounds to 1 for x just below an integer). */
static inline float fract1(float x)
{
    return fminf(x - floorf(x), 0x1.fffffep-1f);
}

/* A source image as it is sampled: per footprint entry (x, y), x in
 * [-1, W], y in [-1, H] (row pitch W + 2), the four bilinear coefficients
 * of texels (x..x+1, y..y+1) with the edge texels repeated (clamp
 * addressing), hpm_bilinear_coeffs of the FP16-rounded texels img, held as
 * floats. */
static void make_coef_plane(const float* img, int W, int H, std::vector<float>* out)
{
    const int pitch = W + 2;
    out->assign((size_t)pitch * (H + 2) * 4, 0.0f);
    auto at = [&](int x, int y) { return img[(size_t)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1)]; };
    for (int y = -1; y <= H; ++y) {
        for (int x = -1; x <= W; ++x) {
            uint16_t c[4];
            hpm_bilinear_coeffs(at(x, y), at(x + 1, y), at(x, y + 1), at(x + 1, y + 1), c);
            float* d = out->data() + 4 * ((size_t)(y + 1) * pitch + (x + 1));
            for (int k = 0; k < 4; ++k) {
                d[k] = hpm_f16_value(c[k]);
            }
        }
    }
}

/* the sample at fractions (a, b) of footprint entry (ix, iy), ix in
 * [-1, W], iy in [-1, H]: hpm_bilinear_eval */
static inline float sample_coef(const float* coef, int W, int ix, int iy, float a, float b)
{
    const float* c = coef + 4 * ((size_t)(iy + 1) * (W + 2) + (ix + 1));
    const float v0 = fmaf(a, c[1], c[0]);
    const float v = fmaf(a, c[3], c[2]);
    return fmaf(b, v, v0);
}

/* tex2D with cudaFilterModeLinear and clamp addressing (HPM.cpp:599-617) at
 * the tap (dx, dy) of a warp; texel centres at integer coordinates (the
 * reference samples at +0.5). */
static inline float sample_tap(const float* coef, int W, int H, const Warp& w, float dx, float dy)
{
    /* the dx products are per window column, the dy terms fused per tap */
    const float u = fmaf(w.ey, dy, w.ex * dx);
    const float inv = tap_inv(u, w, dx, dy);
    float lx = fmaf(fmaf(w.gxy, dy, w.gxx * dx), inv, w.fx0);
    float ly = fmaf(fmaf(w.gyy, dy, w.gyx * dx), inv, w.fy0);
    lx = fminf(fmaxf(lx, -1.0f - (float)w.ix0), (float)(W - w.ix0));
    ly = fminf(fmaxf(ly, -1.0f - (float)w.iy0), (float)(H - w.iy0));
    const float a = fract1(lx);
    const float b = fract1(ly);
    const int ix = w.ix0 + (int)floorf(lx);
    const int iy = w.iy0 + (int)floorf(ly);
    return sample_coef(coef, W, ix, iy, a, b);
}

static inline float texel(const float* img, int W, int H, int x, int y)
{
    return img[(size_t)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1)];
}

/* Reference-only part of ComputeBilateralNCC (HPM.cu:337-369): the window
 * taps, their bilateral weights (ComputeBilateralWeight, HPM.cu:318-323) and
 * the reference sums, all relative to the centre value. */
struct RefPatch {
    int ntap;
    int centre_tap; /* first tap of smallest offset from the centre */
    float dx[64], dy[64];
    float w[64], rr[64], wr[64];
    float c;
    float inv_w;
    float mr;
    float var_r;
};

static void make_ref_patch(const float* img, int W, int H, int px, int py, const hpm_params& P, RefPatch* rp)
{
    const int radius = P.patch_size / 2;
    const float c = texel(img, W, H, px, py);
    float wsum = 0.0f, sr = 0.0f, srr = 0.0f;
    int k = 0;
    for (int i = -radius; i < radius + 1; i += P.radius_increment) {
        for (int j = -radius; j < radius + 1; j += P.radius_increment) {
            const float r = texel(img, W, H, px + i, py + j);
            const float fi = (float)i, fj = (float)j;
            const float sd = sqrtf(fi * fi + fj * fj);
            const float cd = fabsf(r - c);
            const float w = hpm_expf(-sd / (2.0f * P.sigma_spatial * P.sigma_spatial) -
                                     cd / (2.0f * P.sigma_color * P.sigma_color));
            const float rr = r - c;
            const float wr = w * rr;
            rp->dx[k] = fi;
            rp->dy[k] = fj;
            rp->w[k] = w;
            rp->rr[k] = rr;
            rp->wr[k] = wr;
            wsum = wsum + w;
            sr = sr + wr;
            srr = fmaf(wr, rr, srr);
            ++k;
        }
    }
    rp->ntap = k;
    rp->centre_tap = 0;
    for (int t = 1; t < k; ++t) {
        const float dt = rp->dx[t] * rp->dx[t] + rp->dy[t] * rp->dy[t];
        const float db = rp->dx[rp->centre_tap] * rp->dx[rp->centre_tap] + rp->dy[rp->centre_tap] * rp->dy[rp->centre_tap];
        if (dt < db) {
            rp->centre_tap = t;
        }
    }
    rp->c = c;
    rp->inv_w = 1.0f / wsum;
    rp->mr = sr * rp->inv_w;
    rp->var_r = srr * rp->inv_w - rp->mr * rp->mr;
}

struct Ctx {
    const hpm_oracle_problem* pb;
    std::vector<float> src16[HPM_MAX_IMAGES]; /* source images at their FP16-rounded texels */
    std::vector<float> coef[HPM_MAX_IMAGES];  /* source images as sampled (make_coef_plane) */
    int W, H, nsrc;
    ViewRel rel[HPM_MAX_IMAGES];
    RefK rk;
    IterConsts ic;
    float yy_limit; /* unused placeholder for symmetry */
};

/* ComputeBilateralNCC (HPM.cu:325-397) for one source view, FP32 sums;
 * q from plane_q of the hypothesis. */
static float ncc(const Ctx& cx, const RefPatch& rp, int px, int py, int v, const double q[3])
{
    const hpm_oracle_problem* pb = cx.pb;
    const hpm_camera& sc = pb->cams[v];
    const int SW = sc.width, SH = sc.height;
    const float* simg = cx.coef[v].data(); /* as sampled */
    Warp w;
    if (!make_warp(cx.rel[v], q, px, py, SW, SH, (float)(pb->params.patch_size / 2), &w)) {
        return 2.0f;
    }
    float ss_ = 0.0f, sss = 0.0f, srs = 0.0f;
    /* source values are taken relative to the tap nearest the window
     * centre, sampled first: a constant source patch then sums to exactly 0 */
    const int kc = rp.centre_tap;
    const float shift = sample_tap(simg, SW, SH, w, rp.dx[kc], rp.dy[kc]);
    for (int k = 0; k < rp.ntap; ++k) {
        const float raw = k == kc ? shift : sample_tap(simg, SW, SH, w, rp.dx[k], rp.dy[k]);
        const float s = raw - shift;
        const float ws = rp.w[k] * s;
        ss_ = ss_ + ws;
        sss = fmaf(ws, s, sss);
        srs = fmaf(rp.wr[k], s, srs);
    }
    const float ms = ss_ * rp.inv_w;
    const float var_s = sss * rp.inv_w - ms * ms;
    const float kMinVar = 1e-5f;
    if (rp.var_r < kMinVar || var_s < kMinVar) {
        return 2.0f;
    }
    const float cov = srs * rp.inv_w - rp.mr * ms;
    const float den = sqrtf(rp.var_r * var_s);
    return fmaxf(0.0f, fminf(2.0f, 1.0f - cov / den));
}

/* Bilinear sample at an absolute source coordinate (x, y), clamped to
 * [-1, W] x [-1, H] with edge texels repeated: the sampling of sample_tap. */
static inline float sample_abs(const float* coef, int W, int H, float x, float y)
{
    x = fminf(fmaxf(x, -1.0f), (float)W);
    y = fminf(fmaxf(y, -1.0f), (float)H);
    const float a = fract1(x), b = fract1(y);
    return sample_coef(coef, W, (int)floorf(x), (int)floorf(y), a, b);
}

/* Accuracy evidence only: ComputeBilateralNCC (HPM.cu:227-397) in the
 * reference's own FP32 arithmetic order: the homography rebuilt in float
 * per call, each tap's source point divided out in float
 * (ComputeCorrespondingPoint), raw window sums E[xy] - E[x]E[y] in float.
 * Sampling as the FP32 path's (FP16 texels, exact FP32 bilinear). */
static double ncc32_reference_order(const hpm_oracle_problem* pb, const Ctx& cx, int px, int py, int v, F4 pl)
{
    const hpm_camera& rc = pb->cams[0];
    const hpm_camera& sc = pb->cams[v];
    const hpm_params& P = pb->params;
    const int SW = sc.width, SH = sc.height;
    const float* rimg = pb->images[0];
    const float* simg = cx.coef[v].data(); /* as sampled */
    float refC[3], srcC[3], Rr[9], Cr[3], tr[3], H[9], tmp[9];
    for (int a = 0; a < 3; ++a) {
        refC[a] = -(rc.R[a] * rc.t[0] + rc.R[3 + a] * rc.t[1] + rc.R[6 + a] * rc.t[2]);
        srcC[a] = -(sc.R[a] * sc.t[0] + sc.R[3 + a] * sc.t[1] + sc.R[6 + a] * sc.t[2]);
    }
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            Rr[3 * r + c] = sc.R[3 * r] * rc.R[3 * c] + sc.R[3 * r + 1] * rc.R[3 * c + 1] + sc.R[3 * r + 2] * rc.R[3 * c + 2];
        }
        Cr[r] = refC[r] - srcC[r];
    }
    for (int r = 0; r < 3; ++r) {
        tr[r] = sc.R[3 * r] * Cr[0] + sc.R[3 * r + 1] * Cr[1] + sc.R[3 * r + 2] * Cr[2];
    }
    const float nn[3] = {pl.x, pl.y, pl.z};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            H[3 * r + c] = Rr[3 * r + c] - tr[r] * nn[c] / pl.w;
        }
    }
    for (int r = 0; r < 3; ++r) {
        tmp[3 * r + 0] = H[3 * r + 0] / rc.K[0];
        tmp[3 * r + 1] = H[3 * r + 1] / rc.K[4];
        tmp[3 * r + 2] = -H[3 * r + 0] * rc.K[2] / rc.K[0] - H[3 * r + 1] * rc.K[5] / rc.K[4] + H[3 * r + 2];
    }
    H[0] = sc.K[0] * tmp[0] + sc.K[2] * tmp[6];
    H[1] = sc.K[0] * tmp[1] + sc.K[2] * tmp[7];
    H[2] = sc.K[0] * tmp[2] + sc.K[2] * tmp[8];
    H[3] = sc.K[4] * tmp[3] + sc.K[5] * tmp[6];
    H[4] = sc.K[4] * tmp[4] + sc.K[5] * tmp[7];
    H[5] = sc.K[4] * tmp[5] + sc.K[5] * tmp[8];
    H[6] = sc.K[8] * tmp[6];
    H[7] = sc.K[8] * tmp[7];
    H[8] = sc.K[8] * tmp[8];
    auto point = [&](int x, int y, float* ox, float* oy) {
        const float X = (float)x, Y = (float)y;
        const float hx = H[0] * X + H[1] * Y + H[2];
        const float hy = H[3] * X + H[4] * Y + H[5];
        const float hz = H[6] * X + H[7] * Y + H[8];
        *ox = hx / hz;
        *oy = hy / hz;
    };
    float cxp, cyp;
    point(px, py, &cxp, &cyp);
    if (cxp >= SW || cxp < 0.0f || cyp >= SH || cyp < 0.0f) {
        return 2.0;
    }
    const int radius = P.patch_size / 2;
    const float c = texel(rimg, rc.width, rc.height, px, py);
    float sw = 0.0f, sr = 0.0f, srr = 0.0f, ss = 0.0f, sss = 0.0f, srs = 0.0f;
    for (int i = -radius; i < radius + 1; i += P.radius_increment) {
        float swr = 0.0f, srw = 0.0f, srrw = 0.0f, ssw = 0.0f, sssw = 0.0f, srsw = 0.0f;
        for (int j = -radius; j < radius + 1; j += P.radius_increment) {
            const float r = texel(rimg, rc.width, rc.height, px + i, py + j);
            float sx, sy;
            point(px + i, py + j, &sx, &sy);
            const float sv = sample_abs(simg, SW, SH, sx, sy);
            const float w = hpm_expf(-sqrtf((float)(i * i + j * j)) / (2.0f * P.sigma_spatial * P.sigma_spatial) -
                                     fabsf(r - c) / (2.0f * P.sigma_color * P.sigma_color));
            srw += w * r;
            srrw += w * r * r;
            ssw += w * sv;
            sssw += w * sv * sv;
            srsw += w * r * sv;
            swr += w;
        }
        sr += srw;
        srr += srrw;
        ss += ssw;
        sss += sssw;
        srs += srsw;
        sw += swr;
    }
    const float inv = 1.0f / sw;
    sr *= inv;
    srr *= inv;
    ss *= inv;
    sss *= inv;
    srs *= inv;
    const float vr = srr - sr * sr, vs = sss - ss * ss;
    if (vr < 1e-5f || vs < 1e-5f) {
        return 2.0;
    }
    return fmaxf(0.0f, fminf(2.0f, 1.0f - (srs - sr * ss) / sqrtf(vr * vs)));
}

/* Accuracy evidence only: this build's FP32 window sums with the warp
 * (homography H = A - b q^T, warped centre, tap gradients) evaluated in
 * FP32 instead of FP64. */
static double ncc_warp32(const Ctx& cx, const RefPatch& rp, int px, int py, int v, const double qd[3])
{
    const hpm_oracle_problem* pb = cx.pb;
    const hpm_camera& sc = pb->cams[v];
    const int SW = sc.width, SH = sc.height;
    const float* simg = cx.coef[v].data(); /* as sampled */
    const ViewRel& vr = cx.rel[v];
    const float q[3] = {(float)qd[0], (float)qd[1], (float)qd[2]};
    float H[9];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            H[3 * r + c] = fmaf(-(float)vr.b[r], q[c], (float)vr.A[3 * r + c]);
        }
    }
    const float X = (float)px, Y = (float)py;
    const float hx = fmaf(H[0], X, fmaf(H[1], Y, H[2]));
    const float hy = fmaf(H[3], X, fmaf(H[4], Y, H[5]));
    const float hz = fmaf(H[6], X, fmaf(H[7], Y, H[8]));
    const float icz = 1.0f / hz;
    const float sx0 = hx * icz, sy0 = hy * icz;
    if (!(sx0 >= 0.0f && sx0 < (float)SW && sy0 >= 0.0f && sy0 < (float)SH)) {
        return 2.0;
    }
    Warp w;
    const float fx = floorf(sx0), fy = floorf(sy0);
    w.ix0 = (int)fx;
    w.iy0 = (int)fy;
    w.fx0 = sx0 - fx;
    w.fy0 = sy0 - fy;
    w.ex = H[6] * icz;
    w.ey = H[7] * icz;
    w.gxx = (H[0] - sx0 * H[6]) * icz;
    w.gxy = (H[1] - sx0 * H[7]) * icz;
    w.gyx = (H[3] - sy0 * H[6]) * icz;
    w.gyy = (H[4] - sy0 * H[7]) * icz;
    set_tiers(&w, (float)(pb->params.patch_size / 2));
    float ss_ = 0.0f, sss = 0.0f, srs = 0.0f;
    const int kc = rp.centre_tap;
    const float shift = sample_tap(simg, SW, SH, w, rp.dx[kc], rp.dy[kc]);
    for (int k = 0; k < rp.ntap; ++k) {
        const float raw = k == kc ? shift : sample_tap(simg, SW, SH, w, rp.dx[k], rp.dy[k]);
        const float s = raw - shift;
        const float ws = rp.w[k] * s;
        ss_ = ss_ + ws;
        sss = fmaf(ws, s, sss);
        srs = fmaf(rp.wr[k], s, srs);
    }
    const float ms = ss_ * rp.inv_w;
    const float var_s = sss * rp.inv_w - ms * ms;
    if (rp.var_r < 1e-5f || var_s < 1e-5f) {
        return 2.0;
    }
    const float cov = srs * rp.inv_w - rp.mr * ms;
    return fmaxf(0.0f, fminf(2.0f, 1.0f - cov / sqrtf(rp.var_r * var_s)));
}

/* ComputeBilateralNCC in double precision, reference formula (raw sums). */
static double ncc64(const hpm_oracle_problem* pb, const Ctx& cx, int px, int py, int v, F4 plane)
{
    const hpm_camera& rc = pb->cams[0];
    const hpm_camera& sc = pb->cams[v];
    const hpm_params& P = pb->params;
    const int RW = rc.width, RH = rc.height;
    const int SW = sc.width, SH = sc.height;
    const float* rimg = pb->images[0];
    const float* simg = cx.src16[v].data(); /* the texels the FP32 path samples */
    /* homography in double from the same float inputs */
    double refC[3], srcC[3], Rr[9], Cr[3], tr[3], h[9], tmp[9], H[9];
    for (int a = 0; a < 3; ++a) {
        refC[a] = -((double)rc.R[a] * rc.t[0] + (double)rc.R[3 + a] * rc.t[1] + (double)rc.R[6 + a] * rc.t[2]);
        srcC[a] = -((double)sc.R[a] * sc.t[0] + (double)sc.R[3 + a] * sc.t[1] + (double)sc.R[6 + a] * sc.t[2]);
    }
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            Rr[3 * r + c] = (double)sc.R[3 * r] * rc.R[3 * c] + (double)sc.R[3 * r + 1] * rc.R[3 * c + 1] +
                            (double)sc.R[3 * r + 2] * rc.R[3 * c + 2];
        }
    }
    for (int a = 0; a < 3; ++a) {
        Cr[a] = refC[a] - srcC[a];
    }
    for (int r = 0; r < 3; ++r) {
        tr[r] = (double)sc.R[3 * r] * Cr[0] + (double)sc.R[3 * r + 1] * Cr[1] + (double)sc.R[3 * r + 2] * Cr[2];
    }
    const double nn[3] = {plane.x, plane.y, plane.z};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            h[3 * r + c] = Rr[3 * r + c] - tr[r] * nn[c] / (double)plane.w;
        }
    }
    for (int r = 0; r < 3; ++r) {
        tmp[3 * r + 0] = h[3 * r + 0] / rc.K[0];
        tmp[3 * r + 1] = h[3 * r + 1] / rc.K[4];
        tmp[3 * r + 2] = -h[3 * r + 0] * rc.K[2] / rc.K[0] - h[3 * r + 1] * rc.K[5] / rc.K[4] + h[3 * r + 2];
    }
    H[0] = sc.K[0] * tmp[0] + sc.K[2] * tmp[6];
    H[1] = sc.K[0] * tmp[1] + sc.K[2] * tmp[7];
    H[2] = sc.K[0] * tmp[2] + sc.K[2] * tmp[8];
    H[3] = sc.K[4] * tmp[3] + sc.K[5] * tmp[6];
    H[4] = sc.K[4] * tmp[4] + sc.K[5] * tmp[7];
    H[5] = sc.K[4] * tmp[5] + sc.K[5] * tmp[8];
    H[6] = sc.K[8] * tmp[6];
    H[7] = sc.K[8] * tmp[7];
    H[8] = sc.K[8] * tmp[8];
    auto wp = [&](double X, double Y, double* sx, double* sy) {
        const double z = H[6] * X + H[7] * Y + H[8];
        *sx = (H[0] * X + H[1] * Y + H[2]) / z;
        *sy = (H[3] * X + H[4] * Y + H[5]) / z;
    };
    auto bil = [&](double x, double y) {
        x = std::min(std::max(x, -1.0), (double)SW);
        y = std::min(std::max(y, -1.0), (double)SH);
        const double fx = floor(x), fy = floor(y);
        const double a = x - fx, b = y - fy;
        const int ix = (int)fx, iy = (int)fy;
        const int x0 = clampi(ix, 0, SW - 1), x1 = clampi(ix + 1, 0, SW - 1);
        const int y0 = clampi(iy, 0, SH - 1), y1 = clampi(iy + 1, 0, SH - 1);
        const double t00 = simg[(size_t)y0 * SW + x0], t10 = simg[(size_t)y0 * SW + x1];
        const double t01 = simg[(size_t)y1 * SW + x0], t11 = simg[(size_t)y1 * SW + x1];
        return (1 - a) * (1 - b) * t00 + a * (1 - b) * t10 + (1 - a) * b * t01 + a * b * t11;
    };
    double cxp, cyp;
    wp(px, py, &cxp, &cyp);
    if (cxp >= SW || cxp < 0.0 || cyp >= SH || cyp < 0.0) {
        return 2.0;
    }
    const int radius = P.patch_size / 2;
    const double c = texel(rimg, RW, RH, px, py);
    double sw = 0, sr = 0, srr = 0, ss = 0, sss = 0, srs = 0;
    for (int i = -radius; i < radius + 1; i += P.radius_increment) {
        for (int j = -radius; j < radius + 1; j += P.radius_increment) {
            const double r = texel(rimg, RW, RH, px + i, py + j);
            double sx, sy;
            wp(px + i, py + j, &sx, &sy);
            const double s = bil(sx, sy);
            const double w = exp(-sqrt((double)(i * i + j * j)) / (2.0 * P.sigma_spatial * P.sigma_spatial) -
                                 fabs(r - c) / (2.0 * P.sigma_color * P.sigma_color));
            sw += w;
            sr += w * r;
            srr += w * r * r;
            ss += w * s;
            sss += w * s * s;
            srs += w * r * s;
        }
    }
    sr /= sw;
    srr /= sw;
    ss /= sw;
    sss /= sw;
    srs /= sw;
    const double vr = srr - sr * sr, vs = sss - ss * ss;
    if (vr < 1e-5 || vs < 1e-5) {
        return 2.0;
    }
    const double cost = 1.0 - (srs - sr * ss) / sqrt(vr * vs);
    return std::max(0.0, std::min(2.0, cost));
}

/* Get3DPoint + GetViewDirection + GetDistance2Origin +
 * ComputeDepthfromPlaneHypothesis (HPM.cu:102-133). */
static inline void point3(const hpm_camera& c, int px, int py, float depth, float X[3])
{
    X[0] = depth * ((float)px - c.K[2]) / c.K[0];
    X[1] = depth * ((float)py - c.K[5]) / c.K[4];
    X[2] = depth;
}

static inline F4 view_direction(const hpm_camera& c, int px, int py, float depth)
{
    float X[3];
    point3(c, px, py, depth, X);
    const float norm = sqrtf(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    F4 v = {X[0] / norm, X[1] / norm, X[2] / norm, 0.0f};
    return v;
}

static inline float dist_to_origin(const hpm_camera& c, int px, int py, float depth, F4 n)
{
    float X[3];
    point3(c, px, py, depth, X);
    return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}

static inline float depth_from_plane(const hpm_camera& c, F4 pl, int px, int py)
{
    return -pl.w * c.K[0] /
           (((float)px - c.K[2]) * pl.x + (c.K[0] / c.K[4]) * ((float)py - c.K[5]) * pl.y + c.K[0] * pl.z);
}

/* NormalizeVec3 (HPM.cu:77-84); rsqrtf replaced by 1/sqrtf. */
static inline void normalize3(F4* v)
{
    const float n2 = v->x * v->x + v->y * v->y + v->z * v->z;
    const float inv = 1.0f / sqrtf(n2);
    v->x *= inv;
    v->y *= inv;
    v->z *= inv;
}

static inline float dot3(F4 a, F4 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

/* GenerateRandomNormal (HPM.cu:135-161). */
static F4 random_normal(const hpm_camera& c, int px, int py, hpm_rng* rng, float depth)
{
    float q1 = 1.0f, q2 = 1.0f, s = 2.0f;
    while (s >= 1.0f) {
        q1 = 2.0f * hpm_rng_uniform(rng) - 1.0f;
        q2 = 2.0f * hpm_rng_uniform(rng) - 1.0f;
        s = q1 * q1 + q2 * q2;
    }
    const float sq = sqrtf(1.0f - s);
    F4 n = {2.0f * q1 * sq, 2.0f * q2 * sq, 1.0f - 2.0f * s, 0.0f};
    const F4 vd = view_direction(c, px, py, depth);
    const float d = n.x * vd.x + n.y * vd.y + n.z * vd.z;
    if (d > 0.0f) {
        n.x = -n.x;
        n.y = -n.y;
        n.z = -n.z;
    }
    normalize3(&n);
    return n;
}

/* GeneratePerturbedNormal (HPM.cu:163-198). */
static F4 perturbed_normal(const hpm_camera& c, int px, int py, F4 normal, hpm_rng* rng, float perturbation)
{
    const F4 vd = view_direction(c, px, py, 1.0f);
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
    F4 np;
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

/* Get3DPointonWorld_cu / ProjectonCamera_cu (HPM.cu:445-481). */
static inline void point_world(float x, float y, float depth, const hpm_camera& c, float out[3])
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

static inline void project(const float X[3], const hpm_camera& c, float* u, float* v, float* d)
{
    float t[3];
    t[0] = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2] + c.t[0];
    t[1] = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2] + c.t[1];
    t[2] = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2] + c.t[2];
    *d = c.K[6] * t[0] + c.K[7] * t[1] + c.K[8] * t[2];
    *u = (c.K[0] * t[0] + c.K[1] * t[1] + c.K[2] * t[2]) / *d;
    *v = (c.K[3] * t[0] + c.K[4] * t[1] + c.K[5] * t[2]) / *d;
}

/* ComputeGeomConsistencyCost (HPM.cu:483-508; ACMM.cu:494-519 caps it at 3). */
static float geom_cost(const Ctx& cx, int v, F4 plane, int px, int py)
{
    const hpm_oracle_problem* pb = cx.pb;
    const float max_cost = pb->params.acmm ? 3.0f : 5.0f;
    const hpm_camera& rc = pb->cams[0];
    const hpm_camera& sc = pb->cams[v];
    const float depth = depth_from_plane(rc, plane, px, py);
    float fw[3];
    point_world((float)px, (float)py, depth, rc, fw);
    float sx, sy, sd;
    project(fw, sc, &sx, &sy, &sd);
    const float cxf = fminf(fmaxf(sx, -1.0f), (float)sc.width);
    const float cyf = fminf(fmaxf(sy, -1.0f), (float)sc.height);
    const int ix = clampi((int)cxf, 0, sc.width - 1);
    const int iy = clampi((int)cyf, 0, sc.height - 1);
    const float src_depth = pb->depths[v][(size_t)iy * sc.width + ix];
    if (src_depth == 0.0f) {
        return max_cost;
    }
    float bw[3];
    point_world(sx, sy, src_depth, sc, bw);
    float bx, by, bd;
    project(bw, rc, &bx, &by, &bd);
    const float dx = (float)px - bx;
    const float dy = (float)py - by;
    return fminf(max_cost, sqrtf(dx * dx + dy * dy));
}

/* sort_small (HPM.cu:3-12) */
static inline void sort_small(float* d, int n)
{
    int j;
    for (int i = 1; i < n; i++) {
        const float tmp = d[i];
        for (j = i; j >= 1 && tmp < d[j - 1]; j--) {
            d[j] = d[j - 1];
        }
        d[j] = tmp;
    }
}

/* FindMinCostIndex / FindMaxCostIndex (HPM.cu:29-53): the last extremum wins. */
static inline int find_min_idx(const float* c, int n)
{
    float m = c[0];
    int idx = 0;
    for (int i = 1; i < n; ++i) {
        if (c[i] <= m) {
            m = c[i];
            idx = i;
        }
    }
    return idx;
}

static inline int find_max_idx(const float* c, int n)
{
    float m = c[0];
    int idx = 0;
    for (int i = 1; i < n; ++i) {
        if (c[i] >= m) {
            m = c[i];
            idx = i;
        }
    }
    return idx;
}

/* One (pixel, plane, source view) cost: the bilateral NCC, or the
 * micro-scene cost hook when a test installed one. */
static inline float eval_cost(const Ctx& cx, const RefPatch& rp, int px, int py, int v, F4 plane, const double q[3])
{
    if (cx.pb->cost_hook != nullptr) {
        const float pl[4] = {plane.x, plane.y, plane.z, plane.w};
        return cx.pb->cost_hook(cx.pb->cost_hook_user, px, py, v, pl);
    }
    return ncc(cx, rp, px, py, v, q);
}

/* ComputeMultiViewCostVector (HPM.cu:438-443); with weights given, views
 * of zero weight are skipped (the caller never reads their terms). */
static void cost_vector(const Ctx& cx, const RefPatch& rp, int px, int py, F4 plane, float* out,
                        const float* weights, uint64_t* nev)
{
    double q[3];
    plane_q(cx.rk, plane, q);
    for (int v = 1; v <= cx.nsrc; ++v) {
        if (weights != nullptr && !(weights[v - 1] > 0.0f)) {
            out[v - 1] = 0.0f;
            continue;
        }
        out[v - 1] = eval_cost(cx, rp, px, py, v, plane, q);
        ++*nev;
    }
}

/* ComputeMultiViewInitialCostandSelectedViews (HPM.cu:399-436). */
static float initial_cost(const Ctx& cx, const RefPatch& rp, int px, int py, F4 plane, uint32_t* sel,
                          uint64_t* nev)
{
    float cv[32], cvc[32];
    int count = 0, valid = 0;
    double q[3];
    plane_q(cx.rk, plane, q);
    for (int v = 1; v <= cx.nsrc; ++v) {
        const float c = eval_cost(cx, rp, px, py, v, plane, q);
        ++*nev;
        cv[v - 1] = c;
        cvc[v - 1] = c;
        ++count;
        if (c < 2.0f) {
            ++valid;
        }
    }
    sort_small(cv, count);
    *sel = 0u;
    const int top_k = std::min(valid, cx.pb->params.top_k);
    if (top_k > 0) {
        float cost = 0.0f;
        for (int i = 0; i < top_k; ++i) {
            cost += cv[i];
        }
        const float thr = cv[top_k - 1];
        for (int i = 0; i < cx.nsrc; ++i) {
            if (cvc[i] <= thr) {
                *sel |= (1u << i);
            }
        }
        return cost / (float)top_k;
    }
    return 2.0f;
}

static void make_ctx(const hpm_oracle_problem* pb, Ctx* cx)
{
    cx->pb = pb;
    cx->W = pb->cams[0].width;
    cx->H = pb->cams[0].height;
    cx->nsrc = pb->num_images - 1;
    for (int v = 1; v < pb->num_images; ++v) {
        cx->rel[v] = make_view_rel(pb->cams[0], pb->cams[v]);
        const size_t n = (size_t)pb->cams[v].width * pb->cams[v].height;
        std::vector<float>& dst = cx->src16[v];
        dst.resize(n);
        const float* src = pb->images[v];
        for (size_t i = 0; i < n; ++i) {
            dst[i] = hpm_round_f16(src[i]);
        }
        make_coef_plane(dst.data(), pb->cams[v].width, pb->cams[v].height, &cx->coef[v]);
    }
    cx->rk = make_ref_k(pb->cams[0]);
    cx->ic = make_iter_consts();
}

/* One output pixel of JBU_cu (HPM.cu:1984-2052; the same window and
 * weights in RunJBU and the upsample branch of ACMM's RandomInitialization,
 * ACMM.cu:532-601 and 1862-1919): the weighted depth (src_nd.w) and the
 * normalised weighted normal (src_nd.xyz), SpatialGauss / RangeGauss in
 * FP32, NormalizeVec3's rsqrtf as 1/sqrtf. */
static void jbu_pixel(const float* guide, int w, int h, const float* src_nd, int sw, int sh, int scale_i, int px,
                      int py, float* depth_out, F4* normal_out)
{
    const float scale = (float)(1.0 * sw / w);
    const float sigmad = 0.50f, sigmar = 25.5f;
    const int win = scale_i * scale_i + 1;
    const int nn = win / 2;
    const float o_y = (float)py * scale;
    const float o_x = (float)px * scale;
    const float ref_pix = texel(guide, w, h, px, py);
    float nf = 0.0f, ctot = 0.0f;
    F4 ntot = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int j = -nn; j <= nn; ++j) {
        int r_y = (int)(o_y + (float)j);
        r_y = (r_y > 0 ? (r_y < sh ? r_y : sh - 1) : 0);
        const int r_ys = py + j;
        for (int i = -nn; i <= nn; ++i) {
            int r_x = (int)(o_x + (float)i);
            r_x = (r_x > 0 ? (r_x < sw ? r_x : sw - 1) : 0);
            const int s_center = r_y * sw + r_x;
            const float src_pix = src_nd[4 * s_center + 3];
            F4 sn = ld4(src_nd, s_center);
            const int r_xs = px + i;
            const float nb_pix = texel(guide, w, h, r_xs, r_ys);
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
    *normal_out = ntot;
}

/* RandomInitialization (HPM.cu:510-558). */
static void init_pixel(const Ctx& cx, int px, int py, float* hyp, float* cost, uint32_t* sel, uint64_t* nev)
{
    const hpm_oracle_problem* pb = cx.pb;
    const hpm_params& P = pb->params;
    const hpm_camera& rc = pb->cams[0];
    const int center = py * cx.W + px;
    hpm_rng rng;
    hpm_rng_init(&rng, P.seed_lo, P.seed_hi, (uint32_t)center, hpm_stream_init(pb->call_index), pb->salt);
    RefPatch rp;
    make_ref_patch(pb->images[0], cx.W, cx.H, px, py, P, &rp);
    if (P.geom_consistency) {
        F4 t = ld4(hyp, center);
        F4 c;
        c.x = rc.R[0] * t.x + rc.R[1] * t.y + rc.R[2] * t.z;
        c.y = rc.R[3] * t.x + rc.R[4] * t.y + rc.R[5] * t.z;
        c.z = rc.R[6] * t.x + rc.R[7] * t.y + rc.R[8] * t.z;
        c.w = t.w;
        const float depth = c.w;
        c.w = dist_to_origin(rc, px, py, depth, c);
        st4(hyp, center, c);
        cost[center] = initial_cost(cx, rp, px, py, c, &sel[center], nev);
    } else if (P.hierarchy) {
        /* ACMM's upsample branch (ACMM.cu:532-601): depth = the JBU of the
         * coarse depth (RunJBU, ACMM.cpp:804-853, through depths.dmb),
         * normal = the JBU of the coarse normals, normalised, then
         * TransformNormal2RefCam and GetDistance2Origin. pre-cost = the cost
         * of the plane (0, 0, 0, depth): the normal there comes from a fresh
         * host array (ACMM.cpp:711-717), taken as zeros. */
        const int scale_i = (int)std::max((float)cx.W / (float)pb->coarse_w, (float)cx.H / (float)pb->coarse_h);
        float depth;
        F4 n;
        jbu_pixel(pb->images[0], cx.W, cx.H, pb->coarse, pb->coarse_w, pb->coarse_h, scale_i, px, py, &depth, &n);
        const F4 z = {0.0f, 0.0f, 0.0f, depth};
        uint32_t m0;
        pb->pre_cost[center] = initial_cost(cx, rp, px, py, z, &m0, nev);
        F4 c;
        c.x = rc.R[0] * n.x + rc.R[1] * n.y + rc.R[2] * n.z;
        c.y = rc.R[3] * n.x + rc.R[4] * n.y + rc.R[5] * n.z;
        c.z = rc.R[6] * n.x + rc.R[7] * n.y + rc.R[8] * n.z;
        c.w = dist_to_origin(rc, px, py, depth, c);
        st4(hyp, center, c);
        cost[center] = initial_cost(cx, rp, px, py, c, &sel[center], nev);
    } else if (P.planar_prior) {
        if (pb->prior_mask[center] > 0u && cost[center] >= 0.1f) {
            const float perturbation = 0.02f;
            const F4 pl = ld4(pb->prior_planes, center);
            float dp = pl.w;
            const float dmin_p = (1 - 3 * perturbation) * dp;
            const float dmax_p = (1 + 3 * perturbation) * dp;
            dp = hpm_rng_uniform(&rng) * (dmax_p - dmin_p) + dmin_p;
            F4 np = perturbed_normal(rc, px, py, pl, &rng, (float)((double)(3 * perturbation) * kPi));
            np.w = dp;
            st4(hyp, center, np);
            cost[center] = initial_cost(cx, rp, px, py, np, &sel[center], nev);
        } else {
            F4 t = ld4(hyp, center);
            const float depth = t.w;
            t.w = dist_to_origin(rc, px, py, depth, t);
            st4(hyp, center, t);
            cost[center] = initial_cost(cx, rp, px, py, t, &sel[center], nev);
        }
    } else {
        /* GenerateRandomPlaneHypothesis (HPM.cu:200-206) */
        const float depth = hpm_rng_uniform(&rng) * (P.depth_max - P.depth_min) + P.depth_min;
        F4 n = random_normal(rc, px, py, &rng, depth);
        n.w = dist_to_origin(rc, px, py, depth, n);
        st4(hyp, center, n);
        cost[center] = initial_cost(cx, rp, px, py, n, &sel[center], nev);
    }
}

/* JudgeExtend (HPM.cu:1036-1059). */
static bool judge_extend(const Ctx& cx, int iter, int k, const float* c, bool flag)
{
    if (!flag) {
        return false;
    }
    const float good = cx.ic.good[iter][k];
    const float bad = 1.2f; /* float bad_threshold = 1.2; (HPM.cu:1042): compared in float */
    int good_sum = 0, bad_sum = 0;
    for (int i = 0; i < cx.nsrc; ++i) {
        if (c[i] < good) {
            ++good_sum;
        }
        if (c[i] > bad) {
            ++bad_sum;
        }
    }
    return !(good_sum >= 1 && bad_sum <= 2);
}

/* Restricted cost with the planar prior (HPM.cu:1667-1671, 1200-1204). */
struct PriorConsts {
    float gamma, tdss, tass, beta;
};

static inline PriorConsts prior_consts(const hpm_params& P)
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

static inline float prior_term(const PriorConsts& k, float depth_diff, F4 prior_plane, F4 plane)
{
    const float ac = dot3(prior_plane, plane);
    const float ad = hpm_acosf(ac);
    return k.gamma + hpm_expf(-depth_diff * depth_diff / k.tdss) * hpm_expf(-ad * ad / k.tass);
}

/* CheckerboardPropagation (HPM.cu:1222-1725) + PlaneHypothesisRefinement
 * (HPM.cu:1143-1220). */
static void propagate_pixel(const Ctx& cx, int px, int py, int iter, int colour, float* hyp, float* cost,
                            uint32_t* sel, uint64_t* nev)
{
    const hpm_oracle_problem* pb = cx.pb;
    const hpm_params& P = pb->params;
    const hpm_camera& rc = pb->cams[0];
    const int W = cx.W, H = cx.H, nsrc = cx.nsrc;
    if (px >= W || py >= H) {
        return;
    }
    const int center = py * W + px;
    hpm_oracle_trace* tr = (pb->trace != nullptr && pb->trace->px == px && pb->trace->py == py) ? pb->trace : nullptr;
    hpm_rng rng;
    hpm_rng_init(&rng, P.seed_lo, P.seed_hi, (uint32_t)center, hpm_stream_update(pb->call_index, iter, colour),
                 pb->salt);
    RefPatch rp;
    make_ref_patch(pb->images[0], W, H, px, py, P, &rp);

    int left_up = center - 5 - 6 * W;
    int right_up = center + 6 - 5 * W;
    int right_down = center + 5 + 6 * W;
    int left_down = center - 6 + 5 * W;
    int left_far = center - 5;
    int right_far = center + 5;
    int up_far = center - W * 5;
    int down_far = center + W * 5;

    float ca[8][32];
    memset(ca, 0, sizeof(ca));
    ca[0][0] = 2.0f; /* float cost_array[8][32] = { 2.0f } (HPM.cu:1244) */
    bool flag[8] = {false, false, false, false, false, false, false, false};
    float cmin;
    int cmp;

    /* up_far (HPM.cu:1251-1272) */
    if (py > 4) {
        flag[1] = true;
        cmin = cost[up_far];
        cmp = up_far;
        for (int i = 1; i < 5; ++i) {
            if (py > 4 + 2 * i) {
                const int pt = up_far - 2 * i * W;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        up_far = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, up_far), ca[1], nullptr, nev);
    }
    /* down_far (HPM.cu:1274-1295) */
    if (py < H - 5) {
        flag[3] = true;
        cmin = cost[down_far];
        cmp = down_far;
        for (int i = 1; i < 5; ++i) {
            if (py < H - 5 - 2 * i) {
                const int pt = down_far + 2 * i * W;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        down_far = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, down_far), ca[3], nullptr, nev);
    }
    /* left_far (HPM.cu:1297-1318) */
    if (px > 4) {
        flag[5] = true;
        cmin = cost[left_far];
        cmp = left_far;
        for (int i = 1; i < 5; ++i) {
            if (px > 4 + 2 * i) {
                const int pt = left_far - 2 * i;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        left_far = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, left_far), ca[5], nullptr, nev);
    }
    /* right_far keeps the maximum (HPM.cu:1320-1341) */
    if (px < W - 5) {
        flag[7] = true;
        cmin = cost[right_far];
        cmp = right_far;
        for (int i = 1; i < 5; ++i) {
            if (px < W - 5 - 2 * i) {
                const int pt = right_far + 2 * i;
                if (cmin < cost[pt]) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        right_far = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, right_far), ca[7], nullptr, nev);
    }
    /* left_up (HPM.cu:1344-1376) */
    if (px > 4 && py > 5) {
        flag[0] = true;
        cmin = cost[left_up];
        cmp = left_up;
        int qx = px - 5, qy = py - 6;
        for (int i = 0; i < 7; i++) {
            if (i % 2 == 0) {
                qx -= 2;
            } else {
                qy -= 2;
            }
            if (qx >= 0 && qx < W && qy < H && qy >= 0) {
                const int pt = qy * W + qx;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        left_up = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, left_up), ca[0], nullptr, nev);
    }
    /* right_up (HPM.cu:1378-1409) */
    if (px < W - 6 && py > 4) {
        flag[2] = true;
        cmin = cost[right_up];
        cmp = right_up;
        int qx = px + 6, qy = py - 5;
        for (int i = 0; i < 7; i++) {
            if (i % 2 == 0) {
                qy -= 2;
            } else {
                qx += 2;
            }
            if (qx >= 0 && qx < W && qy < H && qy >= 0) {
                const int pt = qy * W + qx;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        right_up = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, right_up), ca[2], nullptr, nev);
    }
    /* right_down (HPM.cu:1411-1442) */
    if (px < W - 5 && py < H - 6) {
        flag[4] = true;
        cmin = cost[right_down];
        cmp = right_down;
        int qx = px + 5, qy = py + 6;
        for (int i = 0; i < 7; i++) {
            if (i % 2 == 0) {
                qx += 2;
            } else {
                qy += 2;
            }
            if (qx >= 0 && qx < W && qy < H && qy >= 0) {
                const int pt = qy * W + qx;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            } else {
                break;
            }
        }
        right_down = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, right_down), ca[4], nullptr, nev);
    }
    /* left_down: the scan never breaks and re-evaluates the cost vector
     * after every step (HPM.cu:1444-1471); the final state equals one
     * evaluation at the final position. */
    if (px > 5 && py < H - 5) {
        flag[6] = true;
        cmin = cost[left_down];
        cmp = left_down;
        int qx = px - 6, qy = py + 5;
        for (int i = 0; i < 7; i++) {
            if (i % 2 == 0) {
                qy += 2;
            } else {
                qx -= 2;
            }
            if (qx >= 0 && qx < W && qy < H && qy >= 0) {
                const int pt = qy * W + qx;
                if (cost[pt] < cmin) {
                    cmin = cost[pt];
                    cmp = pt;
                }
            }
        }
        left_down = cmp;
        cost_vector(cx, rp, px, py, ld4(hyp, left_down), ca[6], nullptr, nev);
    }
    int pos[8] = {left_up, up_far, right_up, down_far, right_down, left_far, left_down, right_far};
    if (tr) {
        tr->hit = 1;
        for (int i = 0; i < 8; ++i) {
            tr->flag[i] = flag[i] ? 1 : 0;
            tr->pos_scan[i] = pos[i];
        }
        for (int k = 0; k < 3; ++k) {
            tr->judge[k] = -1;
        }
    }

    /* NESP extension: only orientation 0 is ever enabled (HPM.cu:1473-1543);
     * ExtendedLeftUpPropagation (HPM.cu:780-843). */
    for (int k = 0; k < 3; ++k) {
        const bool ext = judge_extend(cx, iter, k, ca[0], flag[0]);
        if (tr) {
            tr->judge[k] = ext ? 1 : 0;
        }
        if (!ext) {
            break;
        }
        int lu = center - 5 - 6 * W - 4 * 2 * k - 4 * 2 * W * k;
        if (px > 4 + 8 * k && py > 5 + 8 * k) {
            bool symbol = true;
            cmin = cost[lu];
            cmp = lu;
            int qx = px - 5 - 8 * k, qy = py - 6 - 8 * k;
            for (int i = 0; i < 7; i++) {
                if (i % 2 == 0) {
                    qx -= 2;
                } else {
                    qy -= 2;
                }
                if (qx >= 0 && qx < W && qy < H && qy >= 0) {
                    const int pt = qy * W + qx;
                    if (cost[pt] < cmin) {
                        cmin = cost[pt];
                        cmp = pt;
                    }
                } else {
                    break;
                }
            }
            if (cmin > cost[pos[0]]) {
                cmin = cost[pos[0]];
                cmp = pos[0];
                symbol = false;
            }
            lu = cmp;
            /* re-evaluating the hypothesis already at pos[0] reproduces
             * ca[0] bit for bit, so it is skipped */
            if (symbol && lu != pos[0]) {
                cost_vector(cx, rp, px, py, ld4(hyp, lu), ca[0], nullptr, nev);
            }
        } else {
            lu = pos[0]; /* flag[0] is true here */
        }
        pos[0] = lu;
    }

    if (tr) {
        memcpy(tr->pos, pos, sizeof(pos));
        memcpy(tr->ca, ca, sizeof(ca));
    }

    /* Multi-hypothesis joint view selection (HPM.cu:1548-1613) */
    float vw[32], vprior[32], probs[32];
    for (int i = 0; i < 32; ++i) {
        vw[i] = 0.0f;
        vprior[i] = 0.0f;
        probs[i] = 0.0f;
    }
    const int nb[4] = {center - W * 5, center + W * 5, center - 5, center + 5};
    for (int i = 0; i < 4; ++i) {
        if (flag[2 * i + 1]) {
            const uint32_t m = sel[nb[i]];
            for (int j = 0; j < nsrc; ++j) {
                if (((m >> j) & 1u) == 1u) {
                    vprior[j] += 0.9f;
                } else {
                    vprior[j] += 0.1f;
                }
            }
        }
    }
    const float thr = cx.ic.thr[iter];
    for (int i = 0; i < nsrc; i++) {
        float count = 0.0f;
        int count_false = 0;
        float tmpw = 0.0f;
        for (int j = 0; j < 8; j++) {
            const float c = ca[j][i];
            if (c < thr) {
                tmpw += hpm_expf(c * c / (-0.18f));
                count++;
            }
            if (c > 1.2f) {
                count_false++;
            }
        }
        if (count > 2 && count_false < 3) {
            probs[i] = tmpw / count;
        } else if (count_false < 3) {
            probs[i] = cx.ic.fallback[iter];
        }
        probs[i] = probs[i] * vprior[i];
    }
    if (tr) {
        memcpy(tr->vprior, vprior, sizeof(vprior));
        memcpy(tr->pdf, probs, sizeof(probs));
    }
    /* TransformPDFToCDF (HPM.cu:86-100) */
    {
        float sum = 0.0f;
        for (int i = 0; i < nsrc; ++i) {
            sum += probs[i];
        }
        const float inv = 1.0f / sum;
        float cum = 0.0f;
        for (int i = 0; i < nsrc; ++i) {
            cum += probs[i] * inv;
            probs[i] = cum;
        }
    }
    if (tr) {
        memcpy(tr->cdf, probs, sizeof(probs));
    }
    for (int s = 0; s < 15; ++s) {
        const float u = hpm_rng_uniform(&rng) - FLT_EPSILON;
        if (tr) {
            tr->draws[s] = u;
        }
        for (int v = 0; v < nsrc; ++v) {
            if (probs[v] > u) {
                vw[v] += 1.0f;
                break;
            }
        }
    }
    uint32_t tsel = 0u;
    float wn = 0.0f;
    for (int i = 0; i < nsrc; ++i) {
        if (vw[i] > 0) {
            tsel |= (1u << i);
            wn += vw[i];
        }
    }

    if (tr) {
        memcpy(tr->vw, vw, sizeof(vw));
        tr->tsel = tsel;
        tr->wn = wn;
    }

    /* Candidate scoring (HPM.cu:1615-1633) */
    float fc[8];
    for (int i = 0; i < 8; ++i) {
        fc[i] = 0.0f;
        for (int j = 0; j < nsrc; ++j) {
            if (vw[j] > 0) {
                if (P.geom_consistency) {
                    if (flag[i]) {
                        fc[i] += vw[j] * (ca[i][j] + 0.1f * geom_cost(cx, j + 1, ld4(hyp, pos[i]), px, py));
                    } else {
                        fc[i] += vw[j] * (ca[i][j] + 0.1f * 5.0f);
                    }
                } else {
                    fc[i] += vw[j] * ca[i][j];
                }
            }
        }
        fc[i] /= wn;
    }
    const int min_idx = find_min_idx(fc, 8);
    if (tr) {
        memcpy(tr->fc, fc, sizeof(fc));
        tr->min_idx = min_idx;
        tr->max_idx = -1;
        tr->rcn = 0.0f;
        tr->adopted = -1;
        tr->nan_fallback = 0;
        for (int i = 0; i < 8; ++i) {
            tr->rfc[i] = 0.0f;
        }
    }

    /* Re-score the current hypothesis (HPM.cu:1637-1650); the geometric
     * term weighs 0.1 there and 0.2 in ACMM (ACMM.cu:1596, 651) */
    const float gc_weight = P.acmm ? 0.2f : 0.1f;
    float cvn[32];
    const F4 cur = ld4(hyp, center);
    cost_vector(cx, rp, px, py, cur, cvn, vw, nev);
    float cost_now = 0.0f;
    for (int i = 0; i < nsrc; ++i) {
        if (!(vw[i] > 0.0f)) {
            continue; /* the term is +0 */
        }
        if (P.geom_consistency) {
            cost_now += vw[i] * (cvn[i] + gc_weight * geom_cost(cx, i + 1, cur, px, py));
        } else {
            cost_now += vw[i] * cvn[i];
        }
    }
    cost_now /= wn;
    cost[center] = cost_now;
    if (tr) {
        tr->cost_now = cost_now;
    }
    float depth_now = depth_from_plane(rc, cur, px, py);

    /* Adoption (HPM.cu:1652-1722) */
    float restricted_cost = 0.0f;
    const PriorConsts pk = prior_consts(P);
    if (P.planar_prior) {
        const F4 prior_plane = ld4(pb->prior_planes, center);
        const float depth_prior = depth_from_plane(rc, prior_plane, px, py);
        if (pb->prior_mask[center] > 0u) {
            float rfc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 8; i++) {
                if (flag[i]) {
                    const F4 hp = ld4(hyp, pos[i]);
                    const float dn = depth_from_plane(rc, hp, px, py);
                    const float pr = prior_term(pk, dn - depth_prior, prior_plane, hp);
                    rfc[i] = hpm_expf(-fc[i] * fc[i] / pk.beta) * pr;
                }
            }
            const int max_idx = find_max_idx(rfc, 8);
            const float dn = depth_from_plane(rc, cur, px, py);
            const float pr = prior_term(pk, dn - depth_prior, prior_plane, cur);
            const float rcn = hpm_expf(-cost_now * cost_now / pk.beta) * pr;
            if (tr) {
                memcpy(tr->rfc, rfc, sizeof(rfc));
                tr->max_idx = max_idx;
                tr->rcn = rcn;
            }
            if (flag[max_idx]) {
                const F4 hp = ld4(hyp, pos[max_idx]);
                const float db = depth_from_plane(rc, hp, px, py);
                if (db >= P.depth_min && db <= P.depth_max && rfc[max_idx] > rcn) {
                    depth_now = db;
                    st4(hyp, center, hp);
                    cost[center] = fc[max_idx];
                    restricted_cost = rfc[max_idx];
                    sel[center] = tsel;
                    if (tr) {
                        tr->adopted = max_idx;
                    }
                }
            }
        } else if (flag[min_idx]) {
            const F4 hp = ld4(hyp, pos[min_idx]);
            const float db = depth_from_plane(rc, hp, px, py);
            if (db >= P.depth_min && db <= P.depth_max && fc[min_idx] < cost_now) {
                depth_now = db;
                st4(hyp, center, hp);
                cost[center] = fc[min_idx];
                if (tr) {
                    tr->adopted = min_idx;
                }
            }
        }
    }
    if (!P.planar_prior && flag[min_idx]) {
        const F4 hp = ld4(hyp, pos[min_idx]);
        const float db = depth_from_plane(rc, hp, px, py);
        if (db >= P.depth_min && db <= P.depth_max && fc[min_idx] < cost_now) {
            depth_now = db;
            st4(hyp, center, hp);
            cost[center] = fc[min_idx];
            sel[center] = tsel;
            if (tr) {
                tr->adopted = min_idx;
            }
        }
        if (!P.acmm && cost[center] != cost[center]) { /* no such fallback in ACMM (ACMM.cu:1607-1616) */
            depth_now = db;
            st4(hyp, center, hp);
            cost_now = fc[min_idx];
            sel[center] = tsel;
            if (tr) {
                tr->adopted = min_idx;
                tr->nan_fallback = 1;
            }
        }
    }

    /* PlaneHypothesisRefinement (HPM.cu:1143-1220) */
    {
        const float perturbation = 0.02f;
        float depth_prior = 0.0f;
        float depth_rand;
        F4 n_rand;
        const bool use_prior = P.planar_prior && pb->prior_mask[center] > 0u;
        F4 prior_plane = {0.0f, 0.0f, 0.0f, 0.0f};
        const float ds = (P.depth_max - P.depth_min) / 64.0f;
        const float as = (float)(kPi * (double)(5.0f / 180.0f));
        if (use_prior) {
            prior_plane = ld4(pb->prior_planes, center);
            depth_prior = depth_from_plane(rc, prior_plane, px, py);
            depth_rand = hpm_rng_uniform(&rng) * 6 * ds + (depth_prior - 3 * ds);
            n_rand = perturbed_normal(rc, px, py, prior_plane, &rng, as);
        } else {
            depth_rand = hpm_rng_uniform(&rng) * (P.depth_max - P.depth_min) + P.depth_min;
            n_rand = random_normal(rc, px, py, &rng, depth_now);
        }
        const float dmin_p = (1 - perturbation) * depth_now;
        const float dmax_p = (1 + perturbation) * depth_now;
        const float depth_pert = hpm_rng_uniform(&rng) * (dmax_p - dmin_p) + dmin_p;
        const F4 plane_now = ld4(hyp, center);
        const F4 n_pert = perturbed_normal(rc, px, py, plane_now, &rng, (float)((double)perturbation * kPi));
        const float depths[5] = {depth_rand, depth_now, depth_rand, depth_now, depth_pert};
        const F4 normals[5] = {plane_now, n_rand, n_rand, n_pert, plane_now};
        if (tr) {
            tr->restricted_cost = restricted_cost;
        }
        for (int i = 0; i < 5; ++i) {
            float cvt[32];
            F4 t = normals[i];
            t.w = dist_to_origin(rc, px, py, depths[i], t);
            cost_vector(cx, rp, px, py, t, cvt, vw, nev);
            float tc = 0.0f;
            for (int j = 0; j < nsrc; ++j) {
                if (vw[j] > 0) {
                    if (P.geom_consistency) {
                        tc += vw[j] * (cvt[j] + gc_weight * geom_cost(cx, j + 1, t, px, py));
                    } else {
                        tc += vw[j] * cvt[j];
                    }
                }
            }
            tc /= wn;
            const float db = depth_from_plane(rc, t, px, py);
            bool accepted = false;
            float rtc = 0.0f;
            if (use_prior) {
                const float pr = prior_term(pk, depths[i] - depth_prior, prior_plane, t);
                rtc = hpm_expf(-tc * tc / pk.beta) * pr;
                if (db >= P.depth_min && db <= P.depth_max && rtc > restricted_cost) {
                    depth_now = db;
                    st4(hyp, center, t);
                    cost[center] = tc;
                    restricted_cost = rtc;
                    accepted = true;
                }
            } else {
                if (db >= P.depth_min && db <= P.depth_max && tc < cost[center]) {
                    depth_now = db;
                    st4(hyp, center, t);
                    cost[center] = tc;
                    accepted = true;
                }
            }
            if (tr) {
                tr->cand[i][0] = t.x;
                tr->cand[i][1] = t.y;
                tr->cand[i][2] = t.z;
                tr->cand[i][3] = t.w;
                tr->cand_depth[i] = depths[i];
                tr->cand_cost[i] = tc;
                tr->cand_rcost[i] = rtc;
                tr->cand_accepted[i] = accepted ? 1 : 0;
            }
        }
    }

    /* ACMM hierarchy (ACMM.cu:1619-1628): the adopted and refined
     * hypothesis replaces the current one only when its cost beats the
     * pre-cost by 0.1; otherwise the hypothesis stays and the cost is the
     * re-scored one (written before the adoption). The reference refines a
     * local copy that is left uninitialised when nothing is adopted
     * (ACMM.cu:1605); it is taken as the current hypothesis. */
    if (P.hierarchy && !(cost[center] < pb->pre_cost[center] - 0.1f)) {
        st4(hyp, center, cur);
        cost[center] = cost_now;
    }
}

/* GetDepthandNormal (HPM.cu:1751-1764). */
static void finalize_pixel(const Ctx& cx, int px, int py, float* hyp)
{
    const hpm_camera& rc = cx.pb->cams[0];
    const int center = py * cx.W + px;
    F4 h = ld4(hyp, center);
    h.w = depth_from_plane(rc, h, px, py);
    F4 o;
    o.x = rc.R[0] * h.x + rc.R[3] * h.y + rc.R[6] * h.z;
    o.y = rc.R[1] * h.x + rc.R[4] * h.y + rc.R[7] * h.z;
    o.z = rc.R[2] * h.x + rc.R[5] * h.y + rc.R[8] * h.z;
    o.w = h.w;
    st4(hyp, center, o);
}

/* CheckerboardFilter (HPM.cu:1766-1880). */
static void filter_pixel(const Ctx& cx, int px, int py, float* hyp, const float* cost)
{
    const int W = cx.W, H = cx.H;
    if (px >= W || py >= H) {
        return;
    }
    const int center = py * W + px;
    float f[21];
    int n = 0;
    f[n++] = hyp[4 * center + 3];
    const int left = center - 1, leftleft = center - 3;
    const int up = center - W, upup = center - 3 * W;
    const int down = center + W, downdown = center + 3 * W;
    const int right = center + 1, rightright = center + 3;
    if (cost[center] < 0.001f) {
        return;
    }
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
    sort_small(f, n);
    const int m = n / 2;
    if (n % 2 == 0) {
        hyp[4 * center + 3] = (f[m - 1] + f[m]) / 2;
    } else {
        hyp[4 * center + 3] = f[m];
    }
}

/* Pixels of one colour visited by the 32x16-thread checkerboard launch of
 * HPM.cu:1924-1931 with the row mapping of HPM.cu:1729-1734 (black: x+y
 * even). Rows are 2*yy or 2*yy+1 for yy below the launch's y extent, which
 * decides whether the last row of an odd-height image is visited. */
template <typename F>
static void for_colour(const Ctx& cx, int colour, int nthreads, F&& fn)
{
    const int W = cx.W, H = cx.H;
    const int yy_limit = ((H / 2 + 15) / 16) * 16;
    const int ny = std::min(yy_limit, (H + 1) / 2);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int yy = 0; yy < ny; ++yy) {
        for (int x = 0; x < W; ++x) {
            const int odd = x & 1;
            const int y = 2 * yy + (colour == 0 ? odd : 1 - odd);
            if (y < H) {
                fn(x, y);
            }
        }
    }
}

static int check_problem(const hpm_oracle_problem* pb)
{
    if (pb == nullptr || pb->num_images < 2 || pb->num_images > HPM_MAX_IMAGES) {
        return -1;
    }
    for (int i = 0; i < pb->num_images; ++i) {
        if (pb->images[i] == nullptr || pb->cams[i].width <= 0 || pb->cams[i].height <= 0) {
            return -1;
        }
    }
    if (pb->params.radius_increment <= 0 || pb->params.patch_size < 1) {
        return -1;
    }
    const int per_axis = 2 * (pb->params.patch_size / 2) / pb->params.radius_increment + 1;
    if (per_axis > 8) { /* RefPatch holds at most 64 taps */
        return -1;
    }
    if (pb->params.geom_consistency) {
        for (int i = 1; i < pb->num_images; ++i) {
            if (pb->depths[i] == nullptr) {
                return -1;
            }
        }
    }
    if (pb->params.planar_prior && (pb->prior_planes == nullptr || pb->prior_mask == nullptr)) {
        return -1;
    }
    if (pb->params.acmm && (pb->params.max_iterations < 1 || pb->params.max_iterations > 3 ||
                            pb->params.planar_prior)) {
        return -1;
    }
    if (pb->params.hierarchy &&
        (!pb->params.acmm || pb->params.geom_consistency || pb->coarse == nullptr || pb->pre_cost == nullptr ||
         pb->coarse_w < 1 || pb->coarse_h < 1 ||
         std::max(pb->cams[0].height / pb->coarse_h, pb->cams[0].width / pb->coarse_w) < 2)) {
        return -1;
    }
    return 0;
}

static int nthreads_or_default(int n)
{
    return n > 0 ? n : omp_get_max_threads();
}

} // namespace

extern "C" int hpm_oracle_init(const hpm_oracle_problem* pb, float* hyp, float* cost, uint32_t* selmask,
                               uint64_t* n_eval, int num_threads)
{
    if (check_problem(pb) != 0) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    const int nt = nthreads_or_default(num_threads);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : total)
    for (int y = 0; y < cx.H; ++y) {
        uint64_t local = 0;
        for (int x = 0; x < cx.W; ++x) {
            init_pixel(cx, x, y, hyp, cost, selmask, &local);
        }
        total += local;
    }
    if (n_eval) {
        *n_eval += total;
    }
    return 0;
}

extern "C" int hpm_oracle_half_iteration(const hpm_oracle_problem* pb, int iter, int colour, float* hyp,
                                         float* cost, uint32_t* selmask, uint64_t* n_eval, int num_threads)
{
    return hpm_oracle_half_iteration_rows(pb, iter, colour, 0, INT32_MAX, hyp, cost, selmask, n_eval, num_threads);
}

extern "C" int hpm_oracle_half_iteration_rows(const hpm_oracle_problem* pb, int iter, int colour, int y0, int y1,
                                              float* hyp, float* cost, uint32_t* selmask, uint64_t* n_eval,
                                              int num_threads)
{
    if (check_problem(pb) != 0 || iter < 0 || iter > 2 || colour < 0 || colour > 1) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    const int nt = nthreads_or_default(num_threads);
    std::vector<uint64_t> counts((size_t)cx.H, 0);
    for_colour(cx, colour, nt, [&](int x, int y) {
        if (y < y0 || y >= y1) {
            return;
        }
        uint64_t local = 0;
        propagate_pixel(cx, x, y, iter, colour, hyp, cost, selmask, &local);
        counts[(size_t)y] += local;
    });
    if (n_eval) {
        for (uint64_t c : counts) {
            *n_eval += c;
        }
    }
    return 0;
}

extern "C" int hpm_oracle_finalize(const hpm_oracle_problem* pb, float* hyp, int num_threads)
{
    if (check_problem(pb) != 0) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    const int nt = nthreads_or_default(num_threads);
#pragma omp parallel for num_threads(nt)
    for (int y = 0; y < cx.H; ++y) {
        for (int x = 0; x < cx.W; ++x) {
            finalize_pixel(cx, x, y, hyp);
        }
    }
    return 0;
}

extern "C" int hpm_oracle_filter(const hpm_oracle_problem* pb, int colour, float* hyp, const float* cost,
                                 int num_threads)
{
    if (check_problem(pb) != 0 || colour < 0 || colour > 1) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    const int nt = nthreads_or_default(num_threads);
    for_colour(cx, colour, nt, [&](int x, int y) { filter_pixel(cx, x, y, hyp, cost); });
    return 0;
}

extern "C" int hpm_oracle_run_patchmatch(const hpm_oracle_problem* pb, float* hyp, float* cost, uint32_t* selmask,
                                         uint64_t* n_eval, int num_threads)
{
    uint64_t total = 0;
    if (hpm_oracle_init(pb, hyp, cost, selmask, &total, num_threads) != 0) {
        return -1;
    }
    /* hard-coded 3 (HPM.cu:1933); ACMM runs max_iterations (ACMM.cu:1836) */
    const int iterations = pb->params.acmm ? pb->params.max_iterations : 3;
    for (int it = 0; it < iterations; ++it) {
        hpm_oracle_half_iteration(pb, it, 0, hyp, cost, selmask, &total, num_threads);
        hpm_oracle_half_iteration(pb, it, 1, hyp, cost, selmask, &total, num_threads);
    }
    hpm_oracle_finalize(pb, hyp, num_threads);
    hpm_oracle_filter(pb, 0, hyp, cost, num_threads);
    hpm_oracle_filter(pb, 1, hyp, cost, num_threads);
    if (n_eval) {
        *n_eval = total;
    }
    return 0;
}

extern "C" int hpm_oracle_cost_vectors(const hpm_oracle_problem* pb, const float* planes, const int32_t* pix,
                                       int npix, int fp64, double* out, int num_threads)
{
    if (check_problem(pb) != 0 || npix < 0) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    const int nt = nthreads_or_default(num_threads);
    const int nsrc = cx.nsrc;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int i = 0; i < npix; ++i) {
        const int px = pix[2 * i], py = pix[2 * i + 1];
        const F4 pl = ld4(planes, i);
        if (fp64 == 1) {
            for (int v = 1; v <= nsrc; ++v) {
                out[(size_t)i * nsrc + v - 1] = ncc64(pb, cx, px, py, v, pl);
            }
        } else if (fp64 == 2) {
            for (int v = 1; v <= nsrc; ++v) {
                out[(size_t)i * nsrc + v - 1] = ncc32_reference_order(pb, cx, px, py, v, pl);
            }
        } else {
            RefPatch rp;
            make_ref_patch(pb->images[0], cx.W, cx.H, px, py, pb->params, &rp);
            double q[3];
            plane_q(cx.rk, pl, q);
            for (int v = 1; v <= nsrc; ++v) {
                out[(size_t)i * nsrc + v - 1] = fp64 == 3 ? ncc_warp32(cx, rp, px, py, v, q) : ncc(cx, rp, px, py, v, q);
            }
        }
    }
    return 0;
}

/* JBU_cu (HPM.cu:1984-2052) with SpatialGauss / RangeGauss (HPM.cu:1971-1981)
 * in FP32; NormalizeVec3's rsqrtf as 1/sqrtf. */
extern "C" int hpm_oracle_jbu(const float* guide, int w, int h, const float* src_nd, int sw, int sh,
                              float* out_depth, float* out_normal, int num_threads)
{
    const int scale_i = std::max(h / sh, w / sw);
    if (scale_i == 1) {
        return 1;
    }
    const int nt = nthreads_or_default(num_threads);
#pragma omp parallel for num_threads(nt)
    for (int py = 0; py < h; ++py) {
        for (int px = 0; px < w; ++px) {
            const int center = py * w + px;
            F4 n;
            jbu_pixel(guide, w, h, src_nd, sw, sh, scale_i, px, py, &out_depth[center], &n);
            out_normal[4 * center + 0] = n.x;
            out_normal[4 * center + 1] = n.y;
            out_normal[4 * center + 2] = n.z;
            out_normal[4 * center + 3] = 0.0f;
        }
    }
    return 0;
}

extern "C" int hpm_oracle_round_f16(const float* x, float* y, int n)
{
    for (int i = 0; i < n; ++i) {
        y[i] = hpm_round_f16(x[i]);
    }
    return 0;
}

extern "C" int hpm_oracle_fmath(int fn, const float* x, float* y, int n)
{
    for (int i = 0; i < n; ++i) {
        switch (fn) {
        case 0: y[i] = hpm_expf(x[i]); break;
        case 1: y[i] = hpm_sinf(x[i]); break;
        case 2: y[i] = hpm_cosf(x[i]); break;
        case 3: y[i] = hpm_acosf(x[i]); break;
        case 5: y[i] = fract1(x[i]); break;
        case 6: {
            const int32_t f = (int32_t)floorf(x[i]);
            memcpy(&y[i], &f, sizeof(f));
            break;
        }
        default: return -1;
        }
    }
    return 0;
}

extern "C" int hpm_oracle_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    hpm_philox4x32_10(ctr, key, out);
    return 0;
}

extern "C" int hpm_oracle_uniforms(uint32_t seed_lo, uint32_t seed_hi, uint32_t pixel, uint32_t stream,
                                   uint32_t salt, float* out, int n)
{
    hpm_rng r;
    hpm_rng_init(&r, seed_lo, seed_hi, pixel, stream, salt);
    for (int i = 0; i < n; ++i) {
        out[i] = hpm_rng_uniform(&r);
    }
    return 0;
}

extern "C" int hpm_oracle_jbu_prior(const float* guide, int w, int h, const float* src_nd, int sw, int sh,
                                    const hpm_camera* ref, float depth_min, float depth_max, float* planes,
                                    uint32_t* mask, int num_threads)
{
    std::vector<float> d((size_t)w * h), n((size_t)w * h * 4);
    const int rc = hpm_oracle_jbu(guide, w, h, src_nd, sw, sh, d.data(), n.data(), num_threads);
    if (rc != 0) {
        return rc;
    }
    const hpm_camera& c = *ref;
    for (int py = 0; py < h; ++py) {
        for (int px = 0; px < w; ++px) {
            const size_t q = (size_t)py * w + px;
            const float dq = d[q];
            F4 pl = {0.0f, 0.0f, 0.0f, 0.0f};
            uint32_t m = 0u;
            if (dq <= depth_max && dq >= depth_min) {
                m = 1u;
                const F4 nw = ld4(n.data(), (int)q);
                pl.x = c.R[0] * nw.x + c.R[1] * nw.y + c.R[2] * nw.z;
                pl.y = c.R[3] * nw.x + c.R[4] * nw.y + c.R[5] * nw.z;
                pl.z = c.R[6] * nw.x + c.R[7] * nw.y + c.R[8] * nw.z;
                pl.w = dist_to_origin(c, px, py, dq, pl);
            }
            mask[q] = m;
            st4(planes, (int)q, pl);
        }
    }
    return 0;
}

/* ---- RunFusion (main.cpp:502-633) ---- */

namespace {

struct P3 {
    float x, y, z;
};

/* Get3DPointonWorld (HPM.cpp:209-233) */
P3 fusion_world_point(int x, int y, float depth, const hpm_camera& c)
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
void fusion_project(const P3& X, const hpm_camera& c, float* u, float* v, float* depth)
{
    const float tx = c.R[0] * X.x + c.R[1] * X.y + c.R[2] * X.z + c.t[0];
    const float ty = c.R[3] * X.x + c.R[4] * X.y + c.R[5] * X.z + c.t[1];
    const float tz = c.R[6] * X.x + c.R[7] * X.y + c.R[8] * X.z + c.t[2];
    *depth = c.K[6] * tx + c.K[7] * ty + c.K[8] * tz;
    *u = (c.K[0] * tx + c.K[1] * ty + c.K[2] * tz) / *depth;
    *v = (c.K[3] * tx + c.K[4] * ty + c.K[5] * tz) / *depth;
}

/* GetAngle (HPM.cpp:258-268) with the shared acos */
float fusion_angle(const float* a, const float* b)
{
    const float dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const float angle = hpm_acosf(dot);
    return angle != angle ? 0.0f : angle;
}

} // namespace

/* RunFusion of HPM-MVS (main.cpp:502-633) or, with acmm, of ACMM+NESP
 * (ACMM+NESP/windows/main.cpp:164-300): there the point, normal and colour
 * are the means of the reference pixel's and the consistent sources' (the
 * normal through cv::Vec3f /= float, a product with the float reciprocal),
 * and a point needs 2 consistent sources. */
static int64_t fusion(bool acmm, int num_views, const hpm_camera* cams, const float* const* depths,
                      const float* const* normals, const uint8_t* const* colors, const int* num_src,
                      const int* src_ids, int max_src, hpm_point* out, int64_t cap)
{
    std::vector<std::vector<uint8_t>> masks((size_t)num_views);
    for (int i = 0; i < num_views; ++i) {
        masks[(size_t)i].assign((size_t)cams[i].width * cams[i].height, 0);
    }
    int64_t n = 0;
    for (int i = 0; i < num_views; ++i) {
        const hpm_camera& rc = cams[i];
        const int cols = rc.width, rows = rc.height;
        const int nn = num_src[i];
        const int* srcs = src_ids + (size_t)i * max_src;
        int ux[32], uy[32]; /* used_list: per view, not per pixel (main.cpp:556) */
        for (int j = 0; j < nn; ++j) {
            ux[j] = -1;
            uy[j] = -1;
        }
        for (int r = 0; r < rows; ++r) {
            for (int c = 0; c < cols; ++c) {
                const size_t p = (size_t)r * cols + c;
                if (masks[(size_t)i][p] == 1) {
                    continue;
                }
                const float ref_depth = depths[i][p];
                const float* ref_normal = normals[i] + 3 * p;
                if (ref_depth <= 0.0f) {
                    continue;
                }
                const P3 X = fusion_world_point(c, r, ref_depth, rc);
                int num_consistent = 0;
                float dyn = 0.0f;
                /* ACMM's sums, from the reference pixel */
                P3 sum_x = X;
                float sum_n[3] = {ref_normal[0], ref_normal[1], ref_normal[2]};
                float sum_c[3] = {(float)colors[i][3 * p], (float)colors[i][3 * p + 1], (float)colors[i][3 * p + 2]};
                for (int j = 0; j < nn; ++j) {
                    const int s = srcs[j];
                    const hpm_camera& sc = cams[s];
                    float u, v, proj_depth;
                    fusion_project(X, sc, &u, &v, &proj_depth);
                    const int sr = (int)(v + 0.5f), sc_ = (int)(u + 0.5f);
                    if (!(sc_ >= 0 && sc_ < sc.width && sr >= 0 && sr < sc.height)) {
                        continue;
                    }
                    const size_t q = (size_t)sr * sc.width + sc_;
                    if (masks[(size_t)s][q] == 1) {
                        continue;
                    }
                    const float src_depth = depths[s][q];
                    if (src_depth <= 0.0f) {
                        continue;
                    }
                    const P3 Y = fusion_world_point(sc_, sr, src_depth, sc);
                    float tu, tv;
                    fusion_project(Y, rc, &tu, &tv, &proj_depth);
                    const double dx = (double)((float)c - tu), dy = (double)((float)r - tv);
                    const float reproj = (float)sqrt(dx * dx + dy * dy);
                    const float rdd = fabsf(proj_depth - ref_depth) / ref_depth;
                    const float angle = fusion_angle(ref_normal, normals[s] + 3 * q);
                    if (reproj < 2.0f && rdd < 0.01f && angle < 0.174533f) {
                        if (acmm) {
                            const float* sn = normals[s] + 3 * q;
                            sum_x.x += Y.x;
                            sum_x.y += Y.y;
                            sum_x.z += Y.z;
                            sum_n[0] = sum_n[0] + sn[0];
                            sum_n[1] = sum_n[1] + sn[1];
                            sum_n[2] = sum_n[2] + sn[2];
                            sum_c[0] += (float)colors[s][3 * q];
                            sum_c[1] += (float)colors[s][3 * q + 1];
                            sum_c[2] += (float)colors[s][3 * q + 2];
                        }
                        ux[j] = sc_;
                        uy[j] = sr;
                        const float index = reproj + 200.0f * rdd + angle * 10.0f;
                        dyn += hpm_expf(-index);
                        ++num_consistent;
                    }
                }
                const bool keep = acmm ? num_consistent >= 2
                                       : (num_consistent >= 1 && (double)dyn > 0.3 * (double)num_consistent);
                if (keep) {
                    if (n < cap) {
                        hpm_point& pt = out[n];
                        if (acmm) {
                            const float k = num_consistent + 1.0f;
                            const float ik = 1.0f / k;
                            pt.x = sum_x.x / k;
                            pt.y = sum_x.y / k;
                            pt.z = sum_x.z / k;
                            pt.nx = sum_n[0] * ik;
                            pt.ny = sum_n[1] * ik;
                            pt.nz = sum_n[2] * ik;
                            pt.r = (uint8_t)(int)(sum_c[0] / k);
                            pt.g = (uint8_t)(int)(sum_c[1] / k);
                            pt.b = (uint8_t)(int)(sum_c[2] / k);
                        } else {
                            pt.x = X.x;
                            pt.y = X.y;
                            pt.z = X.z;
                            pt.nx = ref_normal[0];
                            pt.ny = ref_normal[1];
                            pt.nz = ref_normal[2];
                            pt.r = colors[i][3 * p];
                            pt.g = colors[i][3 * p + 1];
                            pt.b = colors[i][3 * p + 2];
                        }
                        pt.pad_ = 0;
                    }
                    ++n;
                    for (int j = 0; j < nn; ++j) {
                        if (ux[j] == -1) {
                            continue;
                        }
                        const int s = srcs[j];
                        masks[(size_t)s][(size_t)uy[j] * cams[s].width + ux[j]] = 1;
                    }
                }
            }
        }
    }
    return n;
}

extern "C" int64_t hpm_oracle_fusion(int num_views, const hpm_camera* cams, const float* const* depths,
                                     const float* const* normals, const uint8_t* const* colors, const int* num_src,
                                     const int* src_ids, int max_src, hpm_point* out, int64_t cap)
{
    return fusion(false, num_views, cams, depths, normals, colors, num_src, src_ids, max_src, out, cap);
}

extern "C" int64_t hpm_oracle_fusion_acmm(int num_views, const hpm_camera* cams, const float* const* depths,
                                          const float* const* normals, const uint8_t* const* colors,
                                          const int* num_src, const int* src_ids, int max_src, hpm_point* out,
                                          int64_t cap)
{
    return fusion(true, num_views, cams, depths, normals, colors, num_src, src_ids, max_src, out, cap);
}

/* ComputeGeomConsistencyCost for given camera-frame planes at given pixels
 * against source view v (for the tests). */
extern "C" int hpm_oracle_geom_cost(const hpm_oracle_problem* pb, int v, const float* planes, const int32_t* pix,
                                    int npix, float* out)
{
    if (check_problem(pb) != 0 || v < 1 || v >= pb->num_images || pb->depths[v] == nullptr) {
        return -1;
    }
    Ctx cx;
    make_ctx(pb, &cx);
    for (int i = 0; i < npix; ++i) {
        out[i] = geom_cost(cx, v, ld4(planes, i), pix[2 * i], pix[2 * i + 1]);
    }
    return 0;
}
