/*
 * hpm_runtime.cpp — host side of the C-ABI (include/hpm_capi.h): the
 * device-resident image pool, the per-problem handle (the reference's
 * `class HPM`, HPM-MVS/HPM.h:48-112, HPM.cpp:99-133, 468-757, 901-921,
 * 1215-1307) and the RunPatchMatch launch schedule (HPM.cu:1906-1957).
 *
 * Differences from the reference by design:
 *   - the images of a scene are uploaded once into a pool shared by the
 *     problems of that scene; passes exchange depth maps and hypotheses on
 *     the device (the reference re-reads images and .dmb files per pass);
 *   - one HIP stream per handle and one synchronisation per RunPatchMatch
 *     (the reference synchronises the device after every launch);
 *   - no RNG state buffer: Philox counters replace the 48 B/px cuRAND state;
 *   - status codes and a last-error string instead of exit();
 *   - release functions are idempotent (the reference double-frees, Q29).
 */
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "hpm_capi.h"
#include "hpm_device.h"
#include "hpm_half.h"
#include "hpm_fmath.h"
#include "hpm_internal.h"
#include "hpm_io.h"
#include "hpm_unet.h"

int hpm_fail(hpm_handle* h, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) {
        h->err = buf;
    }
    return code;
}


namespace {

template <typename T>
void dfree(T*& p)
{
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
}

/* Camera-only part of the plane-induced homography (ComputeHomography,
 * HPM.cu:227-287): H = Ks (Rrel - trel n^T / w) Kr^-1 = A - b q^T with
 * Ks = [sK0 0 sK2; 0 sK4 sK5; 0 0 sK8], Kr = [fx 0 cx; 0 fy cy; 0 0 1] as the
 * reference reads them; A = Ks Rrel Kr^-1 and b = Ks trel in double. */
void make_view_rel(const hpm_camera& rc, const hpm_camera& sc, hpm_dev_view* v)
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
    const double ifx = 1.0 / (double)rc.K[0], ify = 1.0 / (double)rc.K[4];
    const double cx = (double)rc.K[2], cy = (double)rc.K[5];
    for (int r = 0; r < 3; ++r) {
        v->A[3 * r + 0] = M[3 * r + 0] * ifx;
        v->A[3 * r + 1] = M[3 * r + 1] * ify;
        v->A[3 * r + 2] = M[3 * r + 2] - M[3 * r + 0] * cx * ifx - M[3 * r + 1] * cy * ify;
    }
    v->b[0] = sK0 * trel[0] + sK2 * trel[2];
    v->b[1] = sK4 * trel[1] + sK5 * trel[2];
    v->b[2] = sK8 * trel[2];
}

/* JudgeExtend thresholds (HPM.cu:1040-1041, double) and view-selection
 * thresholds (HPM.cu:1568,1586, float) */
void iter_consts(hpm_dev_problem* pb)
{
    for (int it = 0; it < 3; ++it) {
        for (int k = 0; k < 3; ++k) {
            pb->good[it][k] = (float)(0.8 * exp((double)(-it * it * (3 - k)) / 90.0));
        }
        const float e = hpm_expf((float)(it * it) / (-90.0f));
        pb->thr[it] = (float)(0.8 * (double)e);
        pb->fallback[it] = hpm_expf(pb->thr[it] * pb->thr[it] / (-0.32f));
    }
}

int check_params(hpm_handle* h, const hpm_params& p)
{
    if (p.patch_size != 11 || p.radius_increment != 2) {
        return hpm_fail(h, HPM_EINVAL, "the HIP path implements the 11x11 window at stride 2 (patch_size=%d, "
                                       "radius_increment=%d)", p.patch_size, p.radius_increment);
    }
    if (p.top_k < 1 || p.top_k > HPM_MAX_SOURCES) {
        return hpm_fail(h, HPM_EINVAL, "top_k out of range: %d", p.top_k);
    }
    if (p.acmm && (p.max_iterations < 1 || p.max_iterations > 3)) {
        return hpm_fail(h, HPM_EINVAL, "the ACMM variant runs 1..3 iterations (max_iterations=%d)", p.max_iterations);
    }
    if (p.acmm && p.planar_prior) {
        return hpm_fail(h, HPM_EINVAL, "the ACMM variant has no planar prior");
    }
    if (p.hierarchy && (!p.acmm || p.geom_consistency)) {
        return hpm_fail(h, HPM_EINVAL, "hierarchy needs the ACMM variant and no geometric consistency");
    }
    return HPM_OK;
}

void release_problem(hpm_handle* h)
{
    dfree(h->d_views);
    dfree(h->d_hyp);
    dfree(h->d_cost);
    dfree(h->d_sel);
    dfree(h->d_vcost);
    h->vcost_cap = 0;
    dfree(h->d_prior);
    dfree(h->d_mask);
    dfree(h->d_neval);
    dfree(h->d_jbu_src);
    h->jbu_src_cap = 0;
    dfree(h->d_coarse);
    dfree(h->d_pre_cost);
    h->coarse_cap = 0;
    h->have_coarse = false;
    dfree(h->d_ws);
    h->ws_cap = 0;
    h->alloc_pixels = 0;
    h->have_prior = false;
    h->have_device = false;
}

void drop_pool(hpm_handle* h)
{
    if (h->own_pool && h->pool) {
        hpm_pool_destroy(h->pool);
    }
    h->pool = nullptr;
    h->own_pool = false;
}

/* HPM.cpp:553-559 */
void set_depth_range(hpm_handle* h)
{
    h->params.depth_min = h->cams[0].depth_min * 0.6f;
    h->params.depth_max = h->cams[0].depth_max * 1.2f;
    h->params.num_images = h->num_images;
    h->params.disparity_min = h->cams[0].K[0] * h->params.baseline / h->params.depth_max;
    h->params.disparity_max = h->cams[0].K[0] * h->params.baseline / h->params.depth_min;
}

void fill_problem(const hpm_handle* h, hpm_dev_problem* pb)
{
    memset(pb, 0, sizeof(*pb));
    pb->rk_ifx = 1.0 / (double)h->cams[0].K[0];
    pb->rk_ify = 1.0 / (double)h->cams[0].K[4];
    pb->rk_cx = (double)h->cams[0].K[2];
    pb->rk_cy = (double)h->cams[0].K[5];
    pb->W = h->cams[0].width;
    pb->H = h->cams[0].height;
    pb->nsrc = h->num_images - 1;
    /* y extent of the reference's 32x16 checkerboard launch (HPM.cu:1926) */
    pb->yy_limit = ((pb->H / 2 + 15) / 16) * 16;
    pb->params = h->params;
    pb->call_index = h->call_index;
    pb->salt = (uint32_t)h->ref_id;
    iter_consts(pb);
    pb->images = h->pool->d_images;
    pb->images_pad = h->pool->d_images_pad;
    pb->depths = h->params.multi_geometry ? h->pool->d_depth_geom : h->pool->d_depth_plain;
    pb->views = h->d_views;
    pb->hyp = h->d_hyp;
    pb->cost = h->d_cost;
    pb->sel = h->d_sel;
    pb->vcost = h->d_vcost;
    pb->prior_planes = h->d_prior;
    pb->prior_mask = h->d_mask;
    pb->n_eval = h->d_neval;
    if (h->have_coarse) {
        pb->coarse = h->d_coarse;
        pb->coarse_w = h->coarse_w;
        pb->coarse_h = h->coarse_h;
        pb->coarse_scale = h->coarse_scale;
        pb->pre_cost = h->d_pre_cost;
    }
}

int readback(hpm_handle* h)
{
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    h->h_hyp.resize(np * 4);
    h->h_cost.resize(np);
    HPM_HIP(h, hipMemcpyAsync(h->h_hyp.data(), h->d_hyp, np * 16, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->h_cost.data(), h->d_cost, np * 4, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    h->mirror_filled = true;
    return HPM_OK;
}

int ensure_ready(hpm_handle* h)
{
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    if (h->params.planar_prior && !h->have_prior) {
        return hpm_fail(h, HPM_EINVAL, "planar prior enabled but no prior uploaded");
    }
    if (h->params.hierarchy && !h->have_coarse) {
        return hpm_fail(h, HPM_EINVAL, "hierarchy enabled but no coarser map given (hpm_set_hierarchy_input)");
    }
    return check_params(h, h->params);
}

int attach_pool(hpm_handle* h, hpm_pool* pool, bool own, int ref_view, const int* srcs, int nsrc)
{
    if (nsrc < 1 || nsrc > HPM_MAX_SOURCES) {
        return hpm_fail(h, HPM_EINVAL, "a problem needs 1..%d source views (got %d)", HPM_MAX_SOURCES, nsrc);
    }
    if (ref_view < 0 || ref_view >= pool->n) {
        return hpm_fail(h, HPM_EINVAL, "reference view %d outside the pool", ref_view);
    }
    for (int i = 0; i < nsrc; ++i) {
        if (srcs[i] < 0 || srcs[i] >= pool->n || srcs[i] == ref_view) {
            return hpm_fail(h, HPM_EINVAL, "source view %d invalid for reference %d", srcs[i], ref_view);
        }
    }
    if (h->pool != pool) {
        drop_pool(h);
    }
    h->pool = pool;
    h->own_pool = own;
    h->views.assign(1, ref_view);
    h->cams.assign(1, pool->cams[(size_t)ref_view]);
    for (int i = 0; i < nsrc; ++i) {
        h->views.push_back(srcs[i]);
        h->cams.push_back(pool->cams[(size_t)srcs[i]]);
    }
    h->num_images = nsrc + 1;
    set_depth_range(h);
    h->have_input = true;
    h->have_device = false;
    return HPM_OK;
}

} // namespace

/* ------------------------------------------------------------------ pool */

extern "C" hpm_pool* hpm_pool_create(int device, int num_views, const hpm_camera* cams, const float* const* images,
                                     int max_image_size)
{
    if (num_views < 1 || !cams || !images) {
        return nullptr;
    }
    hpm_pool* p = new hpm_pool();
    p->device = device < 0 ? 0 : device;
    p->n = num_views;
    for (int i = 0; i < num_views; ++i) {
        hpm_camera cam = cams[i];
        const int W = cam.width, H = cam.height;
        if (W <= 0 || H <= 0 || !images[i]) {
            delete p;
            return nullptr;
        }
        std::vector<float> img(images[i], images[i] + (size_t)W * H);
        /* Scale cameras and images (HPM.cpp:524-551) */
        if (max_image_size > 0 && (W > max_image_size || H > max_image_size)) {
            const float fx = (float)max_image_size / W;
            const float fy = (float)max_image_size / H;
            const float f = std::min(fx, fy);
            const int nw = (int)lroundf(W * f);
            const int nh = (int)lroundf(H * f);
            std::vector<float> out((size_t)nw * nh);
            hpm_resize_bilinear(img.data(), W, H, out.data(), nw, nh);
            img.swap(out);
            const float sx = nw / (float)W, sy = nh / (float)H;
            cam.K[0] *= sx;
            cam.K[2] *= sx;
            cam.K[4] *= sy;
            cam.K[5] *= sy;
            cam.width = nw;
            cam.height = nh;
        }
        p->cams.push_back(cam);
        p->images.push_back(std::move(img));
        p->off.push_back(p->total);
        p->total += (int64_t)p->cams.back().width * p->cams.back().height;
        p->pad_off.push_back(p->pad_total);
        p->pad_total += (int64_t)(p->cams.back().width + 2) * (p->cams.back().height + 2);
    }
    p->have_state.assign((size_t)num_views, 0);
    bool fits = p->pad_total < ((int64_t)1 << 29); /* 32-bit byte offsets of 8-byte entries */
    for (const hpm_camera& c : p->cams) {
        /* the tap offset's 24-bit signed multiply takes the row pitch in
         * bytes and the row index */
        fits = fits && (int64_t)(c.width + 2) * 8 < ((int64_t)1 << 23) && c.height + 2 < (1 << 23);
    }
    if (!fits) {
        delete p;
        return nullptr;
    }
    bool ok = hipSetDevice(p->device) == hipSuccess &&
              hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&p->d_images, (size_t)p->total * 4) == hipSuccess &&
              hipMalloc((void**)&p->d_images_pad, (size_t)p->pad_total * 8) == hipSuccess &&
              hipMalloc((void**)&p->d_depth_plain, (size_t)p->total * 4) == hipSuccess &&
              hipMalloc((void**)&p->d_depth_geom, (size_t)p->total * 4) == hipSuccess &&
              hipMalloc((void**)&p->d_state_hyp, (size_t)p->total * 16) == hipSuccess &&
              hipMalloc((void**)&p->d_state_cost, (size_t)p->total * 4) == hipSuccess;
    std::vector<uint32_t> padded;
    std::vector<float> t16;
    const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int i = 0; ok && i < num_views; ++i) {
        ok = hipMemcpyAsync(p->d_images + p->off[i], p->images[i].data(), p->images[i].size() * 4,
                            hipMemcpyHostToDevice, p->stream) == hipSuccess;
        const int W = p->cams[i].width, H = p->cams[i].height;
        const float* img = p->images[i].data();
        /* padded pool of bilinear footprints: entry (x, y), x in [-1, W],
         * y in [-1, H], at (y + 1) * (W + 2) + x + 1, holds the FP16
         * coefficients of the footprint of texels (x..x+1, y..y+1)
         * (hpm_bilinear_coeffs of the FP16-rounded texels), c0 | c1 << 16
         * then c2 | c3 << 16; borders repeat the edge texels (clamp
         * addressing). Rows are built on host threads. */
        const int pitch = W + 2;
        t16.resize((size_t)W * H);
        padded.assign((size_t)pitch * (H + 2) * 2, 0u);
        auto rows = [&](int t) {
            for (int y = t; y < H; y += nthreads) {
                for (int x = 0; x < W; ++x) {
                    t16[(size_t)y * W + x] = hpm_round_f16(img[(size_t)y * W + x]);
                }
            }
        };
        auto entries = [&](int t) {
            auto at = [&](int x, int y) {
                return t16[(size_t)std::min(std::max(y, 0), H - 1) * W + std::min(std::max(x, 0), W - 1)];
            };
            for (int y = -1 + t; y <= H; y += nthreads) {
                uint32_t* dst = padded.data() + (size_t)(y + 1) * pitch * 2;
                for (int x = -1; x <= W; ++x) {
                    uint16_t c[4];
                    hpm_bilinear_coeffs(at(x, y), at(x + 1, y), at(x, y + 1), at(x + 1, y + 1), c);
                    dst[2 * (x + 1)] = (uint32_t)c[0] | ((uint32_t)c[1] << 16);
                    dst[2 * (x + 1) + 1] = (uint32_t)c[2] | ((uint32_t)c[3] << 16);
                }
            }
        };
        {
            std::vector<std::thread> ts;
            for (int t = 1; t < nthreads; ++t) ts.emplace_back(rows, t);
            rows(0);
            for (auto& th : ts) th.join();
        }
        {
            std::vector<std::thread> ts;
            for (int t = 1; t < nthreads; ++t) ts.emplace_back(entries, t);
            entries(0);
            for (auto& th : ts) th.join();
        }
        ok = ok && hipMemcpy(p->d_images_pad + 2 * p->pad_off[i], padded.data(), padded.size() * 4,
                             hipMemcpyHostToDevice) == hipSuccess;
    }
    ok = ok && hipMemsetAsync(p->d_depth_plain, 0, (size_t)p->total * 4, p->stream) == hipSuccess &&
         hipMemsetAsync(p->d_depth_geom, 0, (size_t)p->total * 4, p->stream) == hipSuccess &&
         hipStreamSynchronize(p->stream) == hipSuccess;
    if (!ok) {
        hpm_pool_destroy(p);
        return nullptr;
    }
    return p;
}

extern "C" void hpm_pool_destroy(hpm_pool* p)
{
    if (!p) {
        return;
    }
    (void)hipSetDevice(p->device);
    for (hpm_handle* h : p->spare) {
        hpm_destroy(h);
    }
    p->spare.clear();
    hpm_lprior_destroy(p->lprior);
    p->lprior = nullptr;
    if (p->stream) {
        (void)hipStreamSynchronize(p->stream);
    }
    dfree(p->d_images);
    dfree(p->d_images_pad);
    dfree(p->d_depth_plain);
    dfree(p->d_depth_geom);
    dfree(p->d_state_hyp);
    dfree(p->d_state_cost);
    if (p->stream) {
        (void)hipStreamDestroy(p->stream);
    }
    delete p;
}

extern "C" int hpm_pool_camera(const hpm_pool* p, int view, hpm_camera* out)
{
    if (!p || !out || view < 0 || view >= p->n) {
        return HPM_EINVAL;
    }
    *out = p->cams[(size_t)view];
    return HPM_OK;
}

extern "C" int hpm_pool_set_depth(hpm_pool* p, int view, int which, const float* depth)
{
    if (!p || !depth || view < 0 || view >= p->n || which < 0 || which > 1) {
        return HPM_EINVAL;
    }
    float* dst = (which == 0 ? p->d_depth_plain : p->d_depth_geom) + p->off[(size_t)view];
    const size_t np = p->images[(size_t)view].size();
    if (hipMemcpyAsync(dst, depth, np * 4, hipMemcpyHostToDevice, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess) {
        return HPM_EDEVICE;
    }
    return HPM_OK;
}

extern "C" int hpm_pool_get_depth(hpm_pool* p, int view, int which, float* depth)
{
    if (!p || !depth || view < 0 || view >= p->n || which < 0 || which > 1) {
        return HPM_EINVAL;
    }
    const float* src = (which == 0 ? p->d_depth_plain : p->d_depth_geom) + p->off[(size_t)view];
    const size_t np = p->images[(size_t)view].size();
    if (hipMemcpyAsync(depth, src, np * 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess) {
        return HPM_EDEVICE;
    }
    return HPM_OK;
}

/* --------------------------------------------------------------- handles */

extern "C" hpm_handle* hpm_create(int device)
{
    hpm_handle* h = new hpm_handle();
    hpm_params_default(&h->params);
    memset(&h->stats, 0, sizeof(h->stats));
    h->device = device < 0 ? 0 : device;
    if (hipSetDevice(h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return nullptr;
    }
    for (auto& e : h->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
    }
    return h;
}

extern "C" void hpm_destroy(hpm_handle* h)
{
    if (!h) {
        return;
    }
    (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
    }
    release_problem(h);
    drop_pool(h);
    for (auto& e : h->ev) {
        if (e) {
            (void)hipEventDestroy(e);
        }
    }
    if (h->stream) {
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

extern "C" const char* hpm_last_error(const hpm_handle* h)
{
    return h ? h->err.c_str() : "null handle";
}

extern "C" int hpm_set_params(hpm_handle* h, const hpm_params* p)
{
    if (!h || !p) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    const int rc = check_params(h, *p);
    if (rc != HPM_OK) {
        return rc;
    }
    h->params = *p;
    if (h->have_input) {
        set_depth_range(h);
    }
    return HPM_OK;
}

extern "C" int hpm_get_params(const hpm_handle* h, hpm_params* p)
{
    if (!h || !p) {
        return HPM_EINVAL;
    }
    *p = h->params;
    return HPM_OK;
}

extern "C" int hpm_set_geom_consistency(hpm_handle* h, int multi_geometry)
{
    if (!h) {
        return HPM_EINVAL;
    }
    h->params.geom_consistency = 1;
    h->params.max_iterations = 2; /* HPM-MVS stores it only (RunPatchMatch runs 3, HPM.cu:1933); ACMM runs it */
    if (multi_geometry) {
        h->params.multi_geometry = 1;
    }
    return HPM_OK;
}

extern "C" int hpm_set_planar_prior(hpm_handle* h)
{
    if (!h) {
        return HPM_EINVAL;
    }
    h->params.planar_prior = 1;
    return HPM_OK;
}

int hpm_set_hierarchy_internal(hpm_handle* h, const float* coarse_world4, int coarse_w, int coarse_h,
                               bool on_device)
{
    if (!h || !coarse_world4 || coarse_w < 1 || coarse_h < 1) {
        return hpm_fail(h, HPM_EINVAL, "bad hierarchy input");
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    if (!h->params.acmm) {
        return hpm_fail(h, HPM_EINVAL, "hierarchy needs the ACMM variant (params.acmm)");
    }
    const int W = h->cams[0].width, H = h->cams[0].height;
    /* RunJBU (ACMM.cpp:804-853): window scale max(rows / s_rows, cols /
     * s_cols); at 1 the reference writes no upsampled map */
    const int scale = std::max(H / coarse_h, W / coarse_w);
    if (scale < 2) {
        return hpm_fail(h, HPM_EINVAL, "hierarchy input %dx%d is not coarser than %dx%d by 2 or more", coarse_w,
                        coarse_h, W, H);
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t nc = (size_t)coarse_w * coarse_h;
    const size_t np = (size_t)W * H;
    if (h->coarse_cap < nc || !h->d_pre_cost) {
        dfree(h->d_coarse);
        dfree(h->d_pre_cost);
        h->coarse_cap = 0;
        HPM_HIP(h, hipMalloc((void**)&h->d_coarse, nc * 16));
        HPM_HIP(h, hipMalloc((void**)&h->d_pre_cost, h->alloc_pixels * 4));
        h->coarse_cap = nc;
    }
    HPM_HIP(h, hipMemcpyAsync(h->d_coarse, coarse_world4, nc * 16,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_pre_cost, 0, np * 4, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    h->coarse_w = coarse_w;
    h->coarse_h = coarse_h;
    /* the in-kernel window scale: int(max(width / scaled_cols, height /
     * scaled_rows)) in float (ACMM.cu:536) */
    h->coarse_scale = (int)std::max((float)W / (float)coarse_w, (float)H / (float)coarse_h);
    h->have_coarse = true;
    h->params.hierarchy = 1;
    return HPM_OK;
}

extern "C" int hpm_set_hierarchy_input(hpm_handle* h, const float* coarse_world4, int coarse_w, int coarse_h)
{
    return hpm_set_hierarchy_internal(h, coarse_world4, coarse_w, coarse_h, false);
}

extern "C" int hpm_input_init_pool(hpm_handle* h, hpm_pool* pool, int ref_view, const int* src_views, int num_src)
{
    if (!h || !pool || (num_src > 0 && !src_views)) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    h->ref_id = ref_view;
    return attach_pool(h, pool, false, ref_view, src_views, num_src);
}

extern "C" int hpm_input_init_memory(hpm_handle* h, int ref_image_id, int num_images, const hpm_camera* cams,
                                     const float* const* images, const float* const* depths)
{
    if (!h || !cams || !images) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (num_images < 2 || num_images > HPM_MAX_IMAGES) {
        return hpm_fail(h, HPM_EINVAL, "num_images must be 2..%d (got %d)", HPM_MAX_IMAGES, num_images);
    }
    for (int i = 0; i < num_images; ++i) {
        if (!images[i] || cams[i].width <= 0 || cams[i].height <= 0) {
            return hpm_fail(h, HPM_EINVAL, "image %d missing or empty", i);
        }
    }
    hpm_pool* pool = hpm_pool_create(h->device, num_images, cams, images, h->params.max_image_size);
    if (!pool) {
        return hpm_fail(h, HPM_EDEVICE, "cannot create the device image pool");
    }
    if (depths) {
        for (int i = 0; i < num_images; ++i) {
            if (!depths[i]) {
                hpm_pool_destroy(pool);
                return hpm_fail(h, HPM_EINVAL, "depth map %d missing", i);
            }
            if (pool->cams[i].width != cams[i].width || pool->cams[i].height != cams[i].height) {
                hpm_pool_destroy(pool);
                return hpm_fail(h, HPM_EINVAL, "depth maps need images within max_image_size");
            }
            hpm_pool_set_depth(pool, i, 0, depths[i]);
            hpm_pool_set_depth(pool, i, 1, depths[i]);
        }
    }
    std::vector<int> srcs;
    for (int i = 1; i < num_images; ++i) {
        srcs.push_back(i);
    }
    h->ref_id = ref_image_id;
    const int rc = attach_pool(h, pool, true, 0, srcs.data(), (int)srcs.size());
    if (rc != HPM_OK) {
        hpm_pool_destroy(pool);
        h->pool = nullptr;
        h->own_pool = false;
        return rc;
    }
    h->ref_id = ref_image_id;
    return HPM_OK;
}

extern "C" int hpm_input_init(hpm_handle* h, const char* dense_folder, int ref_image_id, const int* src_image_ids,
                              int num_src)
{
    if (!h || !dense_folder || (num_src > 0 && !src_image_ids)) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (num_src < 1 || num_src > HPM_MAX_SOURCES) {
        return hpm_fail(h, HPM_EINVAL, "num_src must be 1..%d (got %d)", HPM_MAX_SOURCES, num_src);
    }
    std::vector<int> ids;
    ids.push_back(ref_image_id);
    for (int i = 0; i < num_src; ++i) {
        ids.push_back(src_image_ids[i]);
    }
    std::vector<std::vector<float>> imgs;
    std::vector<hpm_camera> cams;
    for (int id : ids) {
        char path[4096];
        int w = 0, hh = 0;
        snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder, id);
        if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, id);
            if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
                return hpm_fail(h, HPM_EIO, "cannot read image %d under %s/images (JPEG or 8-bit PGM expected)", id,
                                dense_folder);
            }
        }
        std::vector<float> img((size_t)w * hh);
        if (hpm_read_image_grey(path, img.data(), &w, &hh) != 0) {
            return hpm_fail(h, HPM_EIO, "cannot decode %s", path);
        }
        hpm_camera cam;
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, id);
        if (hpm_read_camera(path, &cam) != 0) {
            return hpm_fail(h, HPM_EIO, "cannot read camera %s", path);
        }
        cam.width = w;
        cam.height = hh;
        cams.push_back(cam);
        imgs.push_back(std::move(img));
    }
    std::vector<std::vector<float>> deps;
    if (h->params.geom_consistency) {
        /* HPM.cpp:561-588 */
        const char* suffix = h->params.multi_geometry ? "depths_geom.dmb" : "depths.dmb";
        for (size_t i = 0; i < ids.size(); ++i) {
            char path[4096];
            snprintf(path, sizeof(path), "%s/HPM/2333_%08d/%s", dense_folder, ids[i], suffix);
            int w = 0, hh = 0, nb = 0;
            if (hpm_read_dmb(path, nullptr, &w, &hh, &nb) != 0 || nb != 1 || w != cams[i].width ||
                hh != cams[i].height) {
                return hpm_fail(h, HPM_EIO, "cannot read depth map %s", path);
            }
            std::vector<float> d((size_t)w * hh);
            if (hpm_read_dmb(path, d.data(), &w, &hh, &nb) != 0) {
                return hpm_fail(h, HPM_EIO, "cannot read depth map %s", path);
            }
            deps.push_back(std::move(d));
        }
    }
    std::vector<const float*> ip, dp;
    for (auto& v : imgs) ip.push_back(v.data());
    for (auto& v : deps) dp.push_back(v.data());
    return hpm_input_init_memory(h, ref_image_id, (int)ids.size(), cams.data(), ip.data(),
                                 deps.empty() ? nullptr : dp.data());
}

extern "C" int hpm_device_init(hpm_handle* h, const char* dense_folder)
{
    if (!h) {
        return HPM_EINVAL;
    }
    if (!h->have_input) {
        return hpm_fail(h, HPM_EINVAL, "input initialisation must come first");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const int n = h->num_images;
    const int W = h->cams[0].width, H = h->cams[0].height;
    const size_t np = (size_t)W * H;
    std::vector<hpm_dev_view> views((size_t)n);
    for (int i = 0; i < n; ++i) {
        memset(&views[i], 0, sizeof(hpm_dev_view));
        views[i].cam = h->cams[i];
        if (i > 0) {
            make_view_rel(h->cams[0], h->cams[i], &views[i]);
        }
        views[i].img_off = h->pool->off[(size_t)h->views[i]];
        views[i].depth_off = views[i].img_off;
        views[i].pad_off = h->pool->pad_off[(size_t)h->views[i]];
        views[i].pitch = views[i].cam.width + 2;
    }
    if (h->alloc_pixels != np) {
        release_problem(h);
        HPM_HIP(h, hipMalloc((void**)&h->d_hyp, np * 16));
        HPM_HIP(h, hipMalloc((void**)&h->d_cost, np * 4));
        HPM_HIP(h, hipMalloc((void**)&h->d_sel, np * 4));
        HPM_HIP(h, hipMalloc((void**)&h->d_neval, 2 * sizeof(unsigned long long)));
        HPM_HIP(h, hipMalloc((void**)&h->d_views, HPM_MAX_IMAGES * sizeof(hpm_dev_view)));
        h->alloc_pixels = np;
    }
    /* the cost cache: one slot per source view and pixel (at least one
     * view, whose slot carries the unknown mark) */
    const size_t vc_need = np * (size_t)std::max(n - 1, 1);
    if (h->vcost_cap < vc_need) {
        dfree(h->d_vcost);
        h->vcost_cap = 0;
        HPM_HIP(h, hipMalloc((void**)&h->d_vcost, vc_need * 4));
        h->vcost_cap = vc_need;
    }
    HPM_HIP(h, hipMemcpyAsync(h->d_views, views.data(), views.size() * sizeof(hpm_dev_view), hipMemcpyHostToDevice,
                              h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_fill_u32((uint32_t*)h->d_vcost, (int64_t)np, 0xffffffffu, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_hyp, 0, np * 16, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_cost, 0, np * 4, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_sel, 0, np * 4, h->stream));
    /* the host mirrors are filled by the first readback (a pass schedule
     * that keeps its maps on the device never pays for them) */
    h->mirror_filled = false;
    h->have_device = true;
    h->have_prior = false;
    h->have_coarse = false;
    h->call_index = 0;
    if (h->params.geom_consistency && dense_folder) {
        /* reload the previous hypotheses (HPM.cpp:662-699) */
        char base[4096];
        snprintf(base, sizeof(base), "%s/%s/2333_%08d", dense_folder, h->params.acmm ? "ACMM_NESP" : "HPM",
                 h->ref_id);
        const char* suffix = h->params.multi_geometry ? "depths_geom.dmb" : "depths.dmb";
        char path[4200];
        std::vector<float> depth(np), normal(np * 3), cost(np);
        /* the header is read first: a file of another size (an earlier run
         * at another scale) is an error, never a read past the buffers */
        auto reload = [&](const char* name, float* dst, int nb_want) {
            snprintf(path, sizeof(path), "%s/%s", base, name);
            int w = 0, hh = 0, nb = 0;
            return hpm_read_dmb(path, nullptr, &w, &hh, &nb) == 0 && w == W && hh == H && nb == nb_want &&
                   hpm_read_dmb(path, dst, &w, &hh, &nb) == 0;
        };
        if (!reload(suffix, depth.data(), 1) || !reload("normals.dmb", normal.data(), 3) ||
            !reload("costs.dmb", cost.data(), 1)) {
            return hpm_fail(h, HPM_EIO, "cannot reload %s (missing, or not %dx%d)", path, W, H);
        }
        std::vector<float> hyp(np * 4);
        for (size_t i = 0; i < np; ++i) {
            hyp[4 * i + 0] = normal[3 * i + 0];
            hyp[4 * i + 1] = normal[3 * i + 1];
            hyp[4 * i + 2] = normal[3 * i + 2];
            hyp[4 * i + 3] = depth[i];
        }
        return hpm_device_init_state(h, hyp.data(), cost.data());
    }
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_set_stream_index(hpm_handle* h, uint32_t index)
{
    if (!h) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    h->call_index = index;
    return HPM_OK;
}

extern "C" int hpm_device_init_state(hpm_handle* h, const float* hyp_world, const float* cost)
{
    if (!h || !hyp_world || !cost) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    h->h_hyp.resize(np * 4);
    h->h_cost.resize(np);
    memcpy(h->h_hyp.data(), hyp_world, np * 16);
    memcpy(h->h_cost.data(), cost, np * 4);
    h->mirror_filled = true;
    HPM_HIP(h, hipMemcpyAsync(h->d_hyp, h->h_hyp.data(), np * 16, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->d_cost, h->h_cost.data(), np * 4, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return hpm_invalidate_cost_cache(h);
}

int hpm_run_internal(hpm_handle* h, bool want_readback)
{
    int rc = ensure_ready(h);
    if (rc != HPM_OK) {
        return rc;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    hpm_dev_problem pb;
    fill_problem(h, &pb);
    HPM_HIP(h, (hipError_t)hpm_launch_zero_u64(h->d_neval, 2, h->stream));
    HPM_HIP(h, hipEventRecord(h->ev[0], h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_init(&pb, h->stream));
    HPM_HIP(h, hipEventRecord(h->ev[1], h->stream));
    pb.n_eval = h->d_neval + 1; /* propagation evaluations counted apart */
    /* hard-coded 3 (HPM.cu:1933); ACMM runs params.max_iterations (ACMM.cu:1836) */
    const int iterations = h->params.acmm ? h->params.max_iterations : 3;
    for (int it = 0; it < iterations; ++it) {
        HPM_HIP(h, (hipError_t)hpm_launch_propagate(&pb, it, 0, h->stream));
        HPM_HIP(h, (hipError_t)hpm_launch_propagate(&pb, it, 1, h->stream));
    }
    HPM_HIP(h, hipEventRecord(h->ev[2], h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_finalize(&pb, h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_filter(&pb, 0, h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_filter(&pb, 1, h->stream));
    HPM_HIP(h, hipEventRecord(h->ev[3], h->stream));
    unsigned long long nev[2] = {0ull, 0ull};
    HPM_HIP(h, hipMemcpyAsync(nev, h->d_neval, sizeof(nev), hipMemcpyDeviceToHost, h->stream));
    if (want_readback) {
        rc = readback(h);
        if (rc != HPM_OK) {
            return rc;
        }
    } else {
        HPM_HIP(h, hipStreamSynchronize(h->stream));
    }
    h->stats.n_eval = nev[0] + nev[1];
    h->stats.n_eval_propagate = nev[1];
    h->stats.launches = 4 + 2 * iterations;
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_total, h->ev[0], h->ev[3]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_init, h->ev[0], h->ev[1]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_propagate, h->ev[1], h->ev[2]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_finalize_filter, h->ev[2], h->ev[3]));
    h->call_index += 1;
    return HPM_OK;
}

extern "C" int hpm_run_patchmatch(hpm_handle* h)
{
    if (!h) {
        return HPM_EINVAL;
    }
    return hpm_run_internal(h, true);
}

int hpm_store_to_pool(hpm_handle* h, int which_depth)
{
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    hpm_pool* p = h->pool;
    const int v = h->views[0];
    const int64_t off = p->off[(size_t)v];
    const int64_t np = (int64_t)h->cams[0].width * h->cams[0].height;
    float* depth = (which_depth == 0 ? p->d_depth_plain : p->d_depth_geom) + off;
    HPM_HIP(h, (hipError_t)hpm_launch_store_state(h->d_hyp, h->d_cost, np, p->d_state_hyp + 4 * off,
                                                  p->d_state_cost + off, depth, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    p->have_state[(size_t)v] = 1;
    return HPM_OK;
}

int hpm_load_from_pool(hpm_handle* h)
{
    hpm_pool* p = h->pool;
    const int v = h->views[0];
    if (!h->have_device || !p->have_state[(size_t)v]) {
        return hpm_fail(h, HPM_EINVAL, "no previous state of view %d in the pool", v);
    }
    const int64_t off = p->off[(size_t)v];
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    HPM_HIP(h, hipMemcpyAsync(h->d_hyp, p->d_state_hyp + 4 * off, np * 16, hipMemcpyDeviceToDevice, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->d_cost, p->d_state_cost + off, np * 4, hipMemcpyDeviceToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return hpm_invalidate_cost_cache(h);
}

int hpm_invalidate_cost_cache(hpm_handle* h)
{
    if (!h->d_vcost) {
        return HPM_OK;
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    HPM_HIP(h, (hipError_t)hpm_launch_fill_u32((uint32_t*)h->d_vcost, (int64_t)np, 0xffffffffu, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

static int ensure_prior_buffers(hpm_handle* h)
{
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (!h->d_prior) {
        HPM_HIP(h, hipMalloc((void**)&h->d_prior, np * 16));
        HPM_HIP(h, hipMalloc((void**)&h->d_mask, np * 4));
    }
    return HPM_OK;
}

int hpm_jbu_prior(hpm_handle* h, const float* src_nd, int sw, int sh, float factor)
{
    (void)factor; /* GetDistance2Origin_factor uses the full-resolution K (HPM.cpp:892-899) */
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const int w = h->cams[0].width, hgt = h->cams[0].height;
    const int scale_i = std::max(hgt / sh, w / sw);
    if (scale_i == 1) {
        return 1;
    }
    int rc = ensure_prior_buffers(h);
    if (rc != HPM_OK) {
        return rc;
    }
    const size_t sn = (size_t)sw * sh;
    if (h->jbu_src_cap < sn * 16) {
        /* kept across calls: a free would synchronise the device */
        dfree(h->d_jbu_src);
        h->jbu_src_cap = 0;
        HPM_HIP(h, hipMalloc((void**)&h->d_jbu_src, sn * 16));
        h->jbu_src_cap = sn * 16;
    }
    hipError_t e = hipMemcpyAsync(h->d_jbu_src, src_nd, sn * 16, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        e = (hipError_t)hpm_launch_jbu_prior(h->pool->d_images + h->pool->off[(size_t)h->views[0]], w, hgt,
                                             h->d_jbu_src, sw, sh, scale_i, &h->cams[0], h->params.depth_min,
                                             h->params.depth_max, h->d_prior, h->d_mask, h->stream);
    }
    if (e == hipSuccess) {
        e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) {
        return hpm_fail(h, HPM_EDEVICE, "JBU prior failed: %s", hipGetErrorString(e));
    }
    h->have_prior = true;
    return HPM_OK;
}

extern "C" int hpm_get_run_stats(const hpm_handle* h, hpm_run_stats* out)
{
    if (!h || !out) {
        return HPM_EINVAL;
    }
    *out = h->stats;
    return HPM_OK;
}

extern "C" int hpm_dims(const hpm_handle* h, int* width, int* height)
{
    if (!h || !h->have_input) {
        return HPM_EINVAL;
    }
    if (width) {
        *width = h->cams[0].width;
    }
    if (height) {
        *height = h->cams[0].height;
    }
    return HPM_OK;
}

extern "C" int hpm_get_camera(const hpm_handle* h, int image, hpm_camera* out)
{
    if (!h || !h->have_input || !out || image < 0 || image >= h->num_images) {
        return HPM_EINVAL;
    }
    *out = h->cams[(size_t)image];
    return HPM_OK;
}

extern "C" int hpm_copy_reference_image(const hpm_handle* h, float* out)
{
    if (!h || !h->have_input || !out) {
        return HPM_EINVAL;
    }
    const std::vector<float>& img = h->pool->images[(size_t)h->views[0]];
    memcpy(out, img.data(), img.size() * sizeof(float));
    return HPM_OK;
}

/* pixels the getters cover: the reference view's, once the device state exists */
static size_t mirror_pixels(const hpm_handle* h)
{
    return h->have_device ? (size_t)h->cams[0].width * h->cams[0].height : 0;
}

extern "C" int hpm_get_plane(const hpm_handle* h, int index, float out[4])
{
    if (!h || !out || index < 0 || (size_t)index >= mirror_pixels(h)) {
        return HPM_EINVAL;
    }
    if (!h->mirror_filled) {
        memset(out, 0, 16);
        return HPM_OK;
    }
    memcpy(out, &h->h_hyp[(size_t)index * 4], 16);
    return HPM_OK;
}

extern "C" float hpm_get_cost(const hpm_handle* h, int index)
{
    if (!h || index < 0 || (size_t)index >= mirror_pixels(h)) {
        return hpm_qnan();
    }
    return h->mirror_filled ? h->h_cost[(size_t)index] : 0.0f;
}

extern "C" int hpm_copy_hypotheses(const hpm_handle* h, float* out4)
{
    const size_t np = h ? mirror_pixels(h) : 0;
    if (!h || !out4 || np == 0) {
        return HPM_EINVAL;
    }
    if (!h->mirror_filled) {
        memset(out4, 0, np * 16);
        return HPM_OK;
    }
    memcpy(out4, h->h_hyp.data(), np * 16);
    return HPM_OK;
}

extern "C" int hpm_copy_costs(const hpm_handle* h, float* out)
{
    const size_t np = h ? mirror_pixels(h) : 0;
    if (!h || !out || np == 0) {
        return HPM_EINVAL;
    }
    if (!h->mirror_filled) {
        memset(out, 0, np * 4);
        return HPM_OK;
    }
    memcpy(out, h->h_cost.data(), np * 4);
    return HPM_OK;
}

extern "C" int hpm_depth_range(const hpm_handle* h, float* depth_min, float* depth_max)
{
    if (!h) {
        return HPM_EINVAL;
    }
    if (depth_min) {
        *depth_min = h->params.depth_min;
    }
    if (depth_max) {
        *depth_max = h->params.depth_max;
    }
    return HPM_OK;
}

extern "C" int hpm_upload_prior(hpm_handle* h, const float* planes4, const uint32_t* mask)
{
    if (!h || !planes4 || !mask) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    int rc = ensure_prior_buffers(h);
    if (rc != HPM_OK) {
        return rc;
    }
    /* planes are only defined where mask > 0 (HPM.cpp:910-918) */
    std::vector<float> planes(np * 4, 0.0f);
    for (size_t i = 0; i < np; ++i) {
        if (mask[i] > 0u) {
            memcpy(&planes[4 * i], &planes4[4 * i], 16);
        }
    }
    HPM_HIP(h, hipMemcpyAsync(h->d_prior, planes.data(), np * 16, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->d_mask, mask, np * 4, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    h->have_prior = true;
    return HPM_OK;
}

extern "C" int hpm_release_prior(hpm_handle* h)
{
    if (!h) {
        return HPM_EINVAL;
    }
    (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
    }
    dfree(h->d_prior);
    dfree(h->d_mask);
    h->have_prior = false;
    return HPM_OK;
}

extern "C" int hpm_release(hpm_handle* h)
{
    if (!h) {
        return HPM_EINVAL;
    }
    (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
    }
    release_problem(h);
    drop_pool(h);
    h->have_input = false;
    return HPM_OK;
}

extern "C" int hpm_jbu(hpm_handle* h, const float* guide, int w, int hgt, const float* src_nd, int sw, int sh,
                       float* out_depth, float* out_normal)
{
    if (!h || !guide || !src_nd || !out_depth || !out_normal) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (w <= 0 || hgt <= 0 || sw <= 0 || sh <= 0) {
        return hpm_fail(h, HPM_EINVAL, "bad sizes");
    }
    /* Imagescale (HPM.cpp:1218-1222) */
    const int scale_i = std::max(hgt / sh, w / sw);
    if (scale_i == 1) {
        return 1;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t n = (size_t)w * hgt, sn = (size_t)sw * sh;
    float *d_guide = nullptr, *d_src = nullptr, *d_depth = nullptr, *d_normal = nullptr;
    hipError_t e = hipMalloc((void**)&d_guide, n * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&d_src, sn * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&d_depth, n * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&d_normal, n * 16);
    if (e == hipSuccess) e = hipMemcpyAsync(d_guide, guide, n * 4, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_src, src_nd, sn * 16, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        e = (hipError_t)hpm_launch_jbu(d_guide, w, hgt, d_src, sw, sh, scale_i, d_depth, d_normal, h->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_depth, d_depth, n * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out_normal, d_normal, n * 16, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(d_guide);
    dfree(d_src);
    dfree(d_depth);
    dfree(d_normal);
    if (e != hipSuccess) {
        return hpm_fail(h, HPM_EDEVICE, "joint bilateral upsampling failed: %s", hipGetErrorString(e));
    }
    return HPM_OK;
}

extern "C" int hpm_cost_vectors(hpm_handle* h, const float* planes4, const int32_t* pix_xy, int npix, float* out)
{
    if (!h || !planes4 || !pix_xy || !out || npix < 0) {
        return hpm_fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const int W = h->cams[0].width, H = h->cams[0].height;
    for (int i = 0; i < npix; ++i) {
        if (pix_xy[2 * i] < 0 || pix_xy[2 * i] >= W || pix_xy[2 * i + 1] < 0 || pix_xy[2 * i + 1] >= H) {
            return hpm_fail(h, HPM_EINVAL, "pixel %d outside the reference image", i);
        }
    }
    if (npix == 0) {
        return HPM_OK;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const int nsrc = h->num_images - 1;
    float *d_planes = nullptr, *d_out = nullptr;
    int32_t* d_pix = nullptr;
    hpm_dev_problem pb;
    fill_problem(h, &pb);
    hipError_t e = hipMalloc((void**)&d_planes, (size_t)npix * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&d_pix, (size_t)npix * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d_out, (size_t)npix * nsrc * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(d_planes, planes4, (size_t)npix * 16, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pix, pix_xy, (size_t)npix * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = (hipError_t)hpm_launch_cost_vectors(&pb, d_planes, d_pix, npix, d_out, h->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(out, d_out, (size_t)npix * nsrc * 4, hipMemcpyDeviceToHost, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(d_planes);
    dfree(d_pix);
    dfree(d_out);
    if (e != hipSuccess) {
        return hpm_fail(h, HPM_EDEVICE, "cost vectors failed: %s", hipGetErrorString(e));
    }
    return HPM_OK;
}

extern "C" int hpm_debug_stage(hpm_handle* h, int stage, int iter, int colour)
{
    if (!h) {
        return HPM_EINVAL;
    }
    int rc = ensure_ready(h);
    if (rc != HPM_OK) {
        return rc;
    }
    if (iter < 0 || iter > 2 || colour < 0 || colour > 1) {
        return hpm_fail(h, HPM_EINVAL, "iter must be 0..2 and colour 0..1");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    hpm_dev_problem pb;
    fill_problem(h, &pb);
    pb.n_eval = nullptr;
    switch (stage) {
    case 0: HPM_HIP(h, (hipError_t)hpm_launch_init(&pb, h->stream)); break;
    case 1: HPM_HIP(h, (hipError_t)hpm_launch_propagate(&pb, iter, colour, h->stream)); break;
    case 2: HPM_HIP(h, (hipError_t)hpm_launch_finalize(&pb, h->stream)); break;
    case 3: HPM_HIP(h, (hipError_t)hpm_launch_filter(&pb, colour, h->stream)); break;
    case 4: return readback(h);
    default: return hpm_fail(h, HPM_EINVAL, "unknown stage %d", stage);
    }
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_debug_read_state(hpm_handle* h, float* hyp4, float* cost, uint32_t* selmask)
{
    if (!h || !h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (hyp4) HPM_HIP(h, hipMemcpyAsync(hyp4, h->d_hyp, np * 16, hipMemcpyDeviceToHost, h->stream));
    if (cost) HPM_HIP(h, hipMemcpyAsync(cost, h->d_cost, np * 4, hipMemcpyDeviceToHost, h->stream));
    if (selmask) HPM_HIP(h, hipMemcpyAsync(selmask, h->d_sel, np * 4, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_debug_read_prior(hpm_handle* h, float* planes4, uint32_t* mask)
{
    if (!h || !h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    if (!h->d_prior || !h->d_mask) {
        return hpm_fail(h, HPM_EINVAL, "no prior on the device");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (planes4) HPM_HIP(h, hipMemcpyAsync(planes4, h->d_prior, np * 16, hipMemcpyDeviceToHost, h->stream));
    if (mask) HPM_HIP(h, hipMemcpyAsync(mask, h->d_mask, np * 4, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_debug_write_state(hpm_handle* h, const float* hyp4, const float* cost, const uint32_t* selmask)
{
    if (!h || !h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (hyp4) HPM_HIP(h, hipMemcpyAsync(h->d_hyp, hyp4, np * 16, hipMemcpyHostToDevice, h->stream));
    if (cost) HPM_HIP(h, hipMemcpyAsync(h->d_cost, cost, np * 4, hipMemcpyHostToDevice, h->stream));
    if (selmask) HPM_HIP(h, hipMemcpyAsync(h->d_sel, selmask, np * 4, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return hyp4 ? hpm_invalidate_cost_cache(h) : HPM_OK;
}

extern "C" int hpm_device_selftest(hpm_handle* h, int kind, const void* in, void* out, int n)
{
    if (!h || !in || !out || n < 0 || kind < 0 || kind > 9) {
        return hpm_fail(h, HPM_EINVAL, "bad selftest arguments");
    }
    if (n == 0) {
        return HPM_OK;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t in_bytes = (size_t)n * (kind == 4 ? 24 : (kind == 8 || kind == 9 ? 8 : 4));
    const size_t out_bytes = (size_t)n * (kind == 4 ? 16 : (kind == 9 ? 8 : 4));
    void *d_in = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_in, in_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_out, out_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, in, in_bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = (hipError_t)hpm_launch_selftest(kind, d_in, d_out, n, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) {
        return hpm_fail(h, HPM_EDEVICE, "selftest failed: %s", hipGetErrorString(e));
    }
    return HPM_OK;
}

/* ----------------------------------------------- shared with hpm_batch.cpp */

void hpm_fill_problem(const hpm_handle* h, hpm_dev_problem* pb)
{
    fill_problem(h, pb);
}

int hpm_ensure_ready(hpm_handle* h)
{
    return ensure_ready(h);
}

int hpm_prior_on_device(hpm_handle* h)
{
    const int rc = ensure_prior_buffers(h);
    if (rc != HPM_OK) {
        return rc;
    }
    h->params.planar_prior = 1;
    h->have_prior = true;
    return HPM_OK;
}

int hpm_workspace(hpm_handle* h, size_t bytes, void** ws)
{
    if (h->ws_cap < bytes) {
        dfree(h->d_ws);
        h->ws_cap = 0;
        HPM_HIP(h, hipMalloc(&h->d_ws, bytes));
        h->ws_cap = bytes;
    }
    *ws = h->d_ws;
    return HPM_OK;
}
