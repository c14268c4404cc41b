/*
 * hpm_runtime.cpp — host side of the C-ABI (include/hpm_capi.h): the
 * per-problem handle, its device buffers, the RunPatchMatch launch schedule
 * and the host mirrors. Mirrors `class HPM` (HPM-MVS/HPM.h:48-112,
 * HPM.cpp:99-133, 468-757, 901-921, 1215-1307; HPM.cu:1906-1957).
 *
 * Differences from the reference by design:
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
#include <vector>

#include "hpm_capi.h"
#include "hpm_device.h"
#include "hpm_fmath.h"
#include "hpm_io.h"

struct hpm_handle {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hpm_params params;
    int ref_id = 0;
    int num_images = 0;
    std::vector<hpm_camera> cams;
    std::vector<std::vector<float>> images;
    std::vector<std::vector<float>> depths; /* geometric consistency inputs */
    bool have_input = false;
    bool have_device = false;
    uint32_t call_index = 0;

    /* device */
    float* d_images = nullptr;
    float* d_depths = nullptr;
    hpm_dev_view* d_views = nullptr;
    float* d_hyp = nullptr;
    float* d_cost = nullptr;
    uint32_t* d_sel = nullptr;
    float* d_prior = nullptr;
    uint32_t* d_mask = nullptr;
    unsigned long long* d_neval = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool have_prior = false;

    /* host mirrors */
    std::vector<float> h_hyp;
    std::vector<float> h_cost;
    hpm_run_stats stats;
};

namespace {

int fail(hpm_handle* h, int code, const char* fmt, ...)
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

#define HPM_HIP(h, call)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            return fail((h), HPM_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                        \
        }                                                                                           \
    } while (0)

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
        return fail(h, HPM_EINVAL, "the HIP path implements the 11x11 window at stride 2 (patch_size=%d, "
                                   "radius_increment=%d)", p.patch_size, p.radius_increment);
    }
    if (p.top_k < 1 || p.top_k > HPM_MAX_SOURCES) {
        return fail(h, HPM_EINVAL, "top_k out of range: %d", p.top_k);
    }
    return HPM_OK;
}

/* input checks shared by both input-init forms (HPM.cpp:482-589) */
int finish_input(hpm_handle* h)
{
    const int n = (int)h->images.size();
    if (n < 2 || n > HPM_MAX_IMAGES) {
        return fail(h, HPM_EINVAL, "a problem needs 1 reference and 1..%d sources (got %d images)",
                    HPM_MAX_SOURCES, n);
    }
    /* Scale cameras and images (HPM.cpp:524-551) */
    for (int i = 0; i < n; ++i) {
        const int W = h->cams[i].width, H = h->cams[i].height;
        if (W <= 0 || H <= 0) {
            return fail(h, HPM_EINVAL, "image %d has no size", i);
        }
        if (W <= h->params.max_image_size && H <= h->params.max_image_size) {
            continue;
        }
        const float fx = (float)h->params.max_image_size / W;
        const float fy = (float)h->params.max_image_size / H;
        const float f = std::min(fx, fy);
        const int nw = (int)lroundf(W * f);
        const int nh = (int)lroundf(H * f);
        std::vector<float> out((size_t)nw * nh);
        hpm_resize_bilinear(h->images[i].data(), W, H, out.data(), nw, nh);
        h->images[i].swap(out);
        const float sx = nw / (float)W, sy = nh / (float)H;
        h->cams[i].K[0] *= sx;
        h->cams[i].K[2] *= sx;
        h->cams[i].K[4] *= sy;
        h->cams[i].K[5] *= sy;
        h->cams[i].width = nw;
        h->cams[i].height = nh;
    }
    h->params.depth_min = h->cams[0].depth_min * 0.6f;
    h->params.depth_max = h->cams[0].depth_max * 1.2f;
    h->params.num_images = n;
    h->params.disparity_min = h->cams[0].K[0] * h->params.baseline / h->params.depth_max;
    h->params.disparity_max = h->cams[0].K[0] * h->params.baseline / h->params.depth_min;
    h->num_images = n;
    if (h->params.geom_consistency) {
        if ((int)h->depths.size() != n) {
            return fail(h, HPM_EINVAL, "geometric consistency needs a depth map for every image");
        }
        for (int i = 0; i < n; ++i) {
            if (h->depths[i].size() != (size_t)h->cams[i].width * h->cams[i].height) {
                return fail(h, HPM_EINVAL, "depth map %d does not match its image size", i);
            }
        }
    }
    h->have_input = true;
    return HPM_OK;
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
    pb->images = h->d_images;
    pb->depths = h->d_depths;
    pb->views = h->d_views;
    pb->hyp = h->d_hyp;
    pb->cost = h->d_cost;
    pb->sel = h->d_sel;
    pb->prior_planes = h->d_prior;
    pb->prior_mask = h->d_mask;
    pb->n_eval = h->d_neval;
}

int readback(hpm_handle* h)
{
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    h->h_hyp.resize(np * 4);
    h->h_cost.resize(np);
    HPM_HIP(h, hipMemcpyAsync(h->h_hyp.data(), h->d_hyp, np * 16, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->h_cost.data(), h->d_cost, np * 4, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

void release_device(hpm_handle* h)
{
    dfree(h->d_images);
    dfree(h->d_depths);
    dfree(h->d_views);
    dfree(h->d_hyp);
    dfree(h->d_cost);
    dfree(h->d_sel);
    dfree(h->d_prior);
    dfree(h->d_mask);
    dfree(h->d_neval);
    h->have_prior = false;
    h->have_device = false;
}

} // namespace

extern "C" hpm_handle* hpm_create(int device)
{
    hpm_handle* h = new hpm_handle();
    hpm_params_default(&h->params);
    memset(&h->stats, 0, sizeof(h->stats));
    h->device = device < 0 ? 0 : device;
    if (hipSetDevice(h->device) != hipSuccess) {
        delete h;
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
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
    release_device(h);
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
        return fail(h, HPM_EINVAL, "null argument");
    }
    const int rc = check_params(h, *p);
    if (rc != HPM_OK) {
        return rc;
    }
    h->params = *p;
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
    h->params.max_iterations = 2; /* stored only: RunPatchMatch runs 3 (HPM.cu:1933) */
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

extern "C" int hpm_input_init(hpm_handle* h, const char* dense_folder, int ref_image_id, const int* src_image_ids,
                              int num_src)
{
    if (!h || !dense_folder || (num_src > 0 && !src_image_ids)) {
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (num_src < 1 || num_src > HPM_MAX_SOURCES) {
        return fail(h, HPM_EINVAL, "num_src must be 1..%d (got %d)", HPM_MAX_SOURCES, num_src);
    }
    h->images.clear();
    h->cams.clear();
    h->depths.clear();
    h->ref_id = ref_image_id;
    std::vector<int> ids;
    ids.push_back(ref_image_id);
    for (int i = 0; i < num_src; ++i) {
        ids.push_back(src_image_ids[i]);
    }
    for (int id : ids) {
        char path[4096];
        int w = 0, hh = 0;
        std::vector<float> img;
        snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, id);
        if (hpm_read_pgm_f32(path, nullptr, &w, &hh) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder, id);
            if (hpm_read_pgm_f32(path, nullptr, &w, &hh) != 0) {
                return fail(h, HPM_EIO, "cannot read image %d under %s/images (8-bit PGM expected)", id,
                            dense_folder);
            }
        }
        img.resize((size_t)w * hh);
        if (hpm_read_pgm_f32(path, img.data(), &w, &hh) != 0) {
            return fail(h, HPM_EIO, "cannot decode %s", path);
        }
        hpm_camera cam;
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, id);
        if (hpm_read_camera(path, &cam) != 0) {
            return fail(h, HPM_EIO, "cannot read camera %s", path);
        }
        cam.width = w;
        cam.height = hh;
        h->cams.push_back(cam);
        h->images.push_back(std::move(img));
    }
    if (h->params.geom_consistency) {
        const char* suffix = h->params.multi_geometry ? "depths_geom.dmb" : "depths.dmb";
        for (int id : ids) {
            char path[4096];
            snprintf(path, sizeof(path), "%s/HPM/2333_%08d/%s", dense_folder, id, suffix);
            int w = 0, hh = 0, nb = 0;
            if (hpm_read_dmb(path, nullptr, &w, &hh, &nb) != 0 || nb != 1) {
                return fail(h, HPM_EIO, "cannot read depth map %s", path);
            }
            std::vector<float> d((size_t)w * hh);
            if (hpm_read_dmb(path, d.data(), &w, &hh, &nb) != 0) {
                return fail(h, HPM_EIO, "cannot read depth map %s", path);
            }
            h->depths.push_back(std::move(d));
        }
    }
    return finish_input(h);
}

extern "C" int hpm_input_init_memory(hpm_handle* h, int ref_image_id, int num_images, const hpm_camera* cams,
                                     const float* const* images, const float* const* depths)
{
    if (!h || !cams || !images) {
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (num_images < 2 || num_images > HPM_MAX_IMAGES) {
        return fail(h, HPM_EINVAL, "num_images must be 2..%d (got %d)", HPM_MAX_IMAGES, num_images);
    }
    h->images.clear();
    h->cams.clear();
    h->depths.clear();
    h->ref_id = ref_image_id;
    for (int i = 0; i < num_images; ++i) {
        if (!images[i] || cams[i].width <= 0 || cams[i].height <= 0) {
            return fail(h, HPM_EINVAL, "image %d missing or empty", i);
        }
        const size_t np = (size_t)cams[i].width * cams[i].height;
        h->cams.push_back(cams[i]);
        h->images.emplace_back(images[i], images[i] + np);
        if (depths) {
            if (!depths[i]) {
                return fail(h, HPM_EINVAL, "depth map %d missing", i);
            }
            h->depths.emplace_back(depths[i], depths[i] + np);
        }
    }
    return finish_input(h);
}

extern "C" int hpm_device_init(hpm_handle* h, const char* dense_folder)
{
    if (!h) {
        return HPM_EINVAL;
    }
    if (!h->have_input) {
        return fail(h, HPM_EINVAL, "input initialisation must come first");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    release_device(h);
    const int n = h->num_images;
    const int W = h->cams[0].width, H = h->cams[0].height;
    const size_t np = (size_t)W * H;
    std::vector<hpm_dev_view> views((size_t)n);
    int64_t off = 0, doff = 0;
    for (int i = 0; i < n; ++i) {
        memset(&views[i], 0, sizeof(hpm_dev_view));
        views[i].cam = h->cams[i];
        if (i > 0) {
            make_view_rel(h->cams[0], h->cams[i], &views[i]);
        }
        views[i].img_off = off;
        views[i].depth_off = doff;
        off += (int64_t)h->cams[i].width * h->cams[i].height;
        if (h->params.geom_consistency) {
            doff += (int64_t)h->cams[i].width * h->cams[i].height;
        }
    }
    HPM_HIP(h, hipMalloc((void**)&h->d_images, (size_t)off * sizeof(float)));
    for (int i = 0; i < n; ++i) {
        HPM_HIP(h, hipMemcpyAsync(h->d_images + views[i].img_off, h->images[i].data(),
                                  h->images[i].size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    }
    if (h->params.geom_consistency) {
        HPM_HIP(h, hipMalloc((void**)&h->d_depths, (size_t)doff * sizeof(float)));
        for (int i = 0; i < n; ++i) {
            HPM_HIP(h, hipMemcpyAsync(h->d_depths + views[i].depth_off, h->depths[i].data(),
                                      h->depths[i].size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
        }
    }
    HPM_HIP(h, hipMalloc((void**)&h->d_views, views.size() * sizeof(hpm_dev_view)));
    HPM_HIP(h, hipMemcpyAsync(h->d_views, views.data(), views.size() * sizeof(hpm_dev_view), hipMemcpyHostToDevice,
                              h->stream));
    HPM_HIP(h, hipMalloc((void**)&h->d_hyp, np * 16));
    HPM_HIP(h, hipMalloc((void**)&h->d_cost, np * 4));
    HPM_HIP(h, hipMalloc((void**)&h->d_sel, np * 4));
    HPM_HIP(h, hipMalloc((void**)&h->d_neval, 2 * sizeof(unsigned long long)));
    HPM_HIP(h, hipMemsetAsync(h->d_hyp, 0, np * 16, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_cost, 0, np * 4, h->stream));
    HPM_HIP(h, hipMemsetAsync(h->d_sel, 0, np * 4, h->stream));
    h->h_hyp.assign(np * 4, 0.0f);
    h->h_cost.assign(np, 0.0f);
    h->have_device = true;
    h->call_index = 0;
    if (h->params.geom_consistency && dense_folder) {
        /* reload the previous hypotheses (HPM.cpp:662-699) */
        char base[4096];
        snprintf(base, sizeof(base), "%s/HPM/2333_%08d", dense_folder, h->ref_id);
        const char* suffix = h->params.multi_geometry ? "depths_geom.dmb" : "depths.dmb";
        char path[4200];
        int w = 0, hh = 0, nb = 0;
        std::vector<float> depth(np), normal(np * 3), cost(np);
        snprintf(path, sizeof(path), "%s/%s", base, suffix);
        if (hpm_read_dmb(path, depth.data(), &w, &hh, &nb) != 0 || w != W || hh != H || nb != 1) {
            return fail(h, HPM_EIO, "cannot reload %s", path);
        }
        snprintf(path, sizeof(path), "%s/normals.dmb", base);
        if (hpm_read_dmb(path, normal.data(), &w, &hh, &nb) != 0 || w != W || hh != H || nb != 3) {
            return fail(h, HPM_EIO, "cannot reload %s", path);
        }
        snprintf(path, sizeof(path), "%s/costs.dmb", base);
        if (hpm_read_dmb(path, cost.data(), &w, &hh, &nb) != 0 || w != W || hh != H || nb != 1) {
            return fail(h, HPM_EIO, "cannot reload %s", path);
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

extern "C" int hpm_device_init_state(hpm_handle* h, const float* hyp_world, const float* cost)
{
    if (!h || !hyp_world || !cost) {
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    memcpy(h->h_hyp.data(), hyp_world, np * 16);
    memcpy(h->h_cost.data(), cost, np * 4);
    HPM_HIP(h, hipMemcpyAsync(h->d_hyp, h->h_hyp.data(), np * 16, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipMemcpyAsync(h->d_cost, h->h_cost.data(), np * 4, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

static int ensure_ready(hpm_handle* h)
{
    if (!h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    if (h->params.planar_prior && !h->have_prior) {
        return fail(h, HPM_EINVAL, "planar prior enabled but no prior uploaded");
    }
    if (h->params.geom_consistency && !h->d_depths) {
        return fail(h, HPM_EINVAL, "geometric consistency enabled after device initialisation");
    }
    return check_params(h, h->params);
}

extern "C" int hpm_run_patchmatch(hpm_handle* h)
{
    if (!h) {
        return HPM_EINVAL;
    }
    int rc = ensure_ready(h);
    if (rc != HPM_OK) {
        return rc;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    hpm_dev_problem pb;
    fill_problem(h, &pb);
    HPM_HIP(h, hipMemsetAsync(h->d_neval, 0, 2 * sizeof(unsigned long long), h->stream));
    HPM_HIP(h, hipEventRecord(h->ev[0], h->stream));
    HPM_HIP(h, (hipError_t)hpm_launch_init(&pb, h->stream));
    HPM_HIP(h, hipEventRecord(h->ev[1], h->stream));
    pb.n_eval = h->d_neval + 1; /* propagation evaluations counted apart */
    for (int it = 0; it < 3; ++it) { /* hard-coded 3 (HPM.cu:1933) */
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
    rc = readback(h);
    if (rc != HPM_OK) {
        return rc;
    }
    h->stats.n_eval = nev[0] + nev[1];
    h->stats.n_eval_propagate = nev[1];
    h->stats.launches = 10;
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_total, h->ev[0], h->ev[3]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_init, h->ev[0], h->ev[1]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_propagate, h->ev[1], h->ev[2]));
    HPM_HIP(h, hipEventElapsedTime(&h->stats.ms_finalize_filter, h->ev[2], h->ev[3]));
    h->call_index += 1;
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

extern "C" int hpm_copy_reference_image(const hpm_handle* h, float* out)
{
    if (!h || !h->have_input || !out) {
        return HPM_EINVAL;
    }
    memcpy(out, h->images[0].data(), h->images[0].size() * sizeof(float));
    return HPM_OK;
}

extern "C" int hpm_get_plane(const hpm_handle* h, int index, float out[4])
{
    if (!h || !out || index < 0 || (size_t)index * 4 >= h->h_hyp.size()) {
        return HPM_EINVAL;
    }
    memcpy(out, &h->h_hyp[(size_t)index * 4], 16);
    return HPM_OK;
}

extern "C" float hpm_get_cost(const hpm_handle* h, int index)
{
    if (!h || index < 0 || (size_t)index >= h->h_cost.size()) {
        return hpm_qnan();
    }
    return h->h_cost[(size_t)index];
}

extern "C" int hpm_copy_hypotheses(const hpm_handle* h, float* out4)
{
    if (!h || !out4 || h->h_hyp.empty()) {
        return HPM_EINVAL;
    }
    memcpy(out4, h->h_hyp.data(), h->h_hyp.size() * sizeof(float));
    return HPM_OK;
}

extern "C" int hpm_copy_costs(const hpm_handle* h, float* out)
{
    if (!h || !out || h->h_cost.empty()) {
        return HPM_EINVAL;
    }
    memcpy(out, h->h_cost.data(), h->h_cost.size() * sizeof(float));
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
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (!h->d_prior) {
        HPM_HIP(h, hipMalloc((void**)&h->d_prior, np * 16));
        HPM_HIP(h, hipMalloc((void**)&h->d_mask, np * 4));
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
    release_device(h);
    h->images.clear();
    h->depths.clear();
    h->have_input = false;
    return HPM_OK;
}

extern "C" int hpm_jbu(hpm_handle* h, const float* guide, int w, int hgt, const float* src_nd, int sw, int sh,
                       float* out_depth, float* out_normal)
{
    if (!h || !guide || !src_nd || !out_depth || !out_normal) {
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (w <= 0 || hgt <= 0 || sw <= 0 || sh <= 0) {
        return fail(h, HPM_EINVAL, "bad sizes");
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
        return fail(h, HPM_EDEVICE, "joint bilateral upsampling failed: %s", hipGetErrorString(e));
    }
    return HPM_OK;
}

extern "C" int hpm_cost_vectors(hpm_handle* h, const float* planes4, const int32_t* pix_xy, int npix, float* out)
{
    if (!h || !planes4 || !pix_xy || !out || npix < 0) {
        return fail(h, HPM_EINVAL, "null argument");
    }
    if (!h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const int W = h->cams[0].width, H = h->cams[0].height;
    for (int i = 0; i < npix; ++i) {
        if (pix_xy[2 * i] < 0 || pix_xy[2 * i] >= W || pix_xy[2 * i + 1] < 0 || pix_xy[2 * i + 1] >= H) {
            return fail(h, HPM_EINVAL, "pixel %d outside the reference image", i);
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
        return fail(h, HPM_EDEVICE, "cost vectors failed: %s", hipGetErrorString(e));
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
        return fail(h, HPM_EINVAL, "iter must be 0..2 and colour 0..1");
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
    default: return fail(h, HPM_EINVAL, "unknown stage %d", stage);
    }
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_debug_read_state(hpm_handle* h, float* hyp4, float* cost, uint32_t* selmask)
{
    if (!h || !h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (hyp4) HPM_HIP(h, hipMemcpyAsync(hyp4, h->d_hyp, np * 16, hipMemcpyDeviceToHost, h->stream));
    if (cost) HPM_HIP(h, hipMemcpyAsync(cost, h->d_cost, np * 4, hipMemcpyDeviceToHost, h->stream));
    if (selmask) HPM_HIP(h, hipMemcpyAsync(selmask, h->d_sel, np * 4, hipMemcpyDeviceToHost, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_debug_write_state(hpm_handle* h, const float* hyp4, const float* cost, const uint32_t* selmask)
{
    if (!h || !h->have_device) {
        return fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
    if (hyp4) HPM_HIP(h, hipMemcpyAsync(h->d_hyp, hyp4, np * 16, hipMemcpyHostToDevice, h->stream));
    if (cost) HPM_HIP(h, hipMemcpyAsync(h->d_cost, cost, np * 4, hipMemcpyHostToDevice, h->stream));
    if (selmask) HPM_HIP(h, hipMemcpyAsync(h->d_sel, selmask, np * 4, hipMemcpyHostToDevice, h->stream));
    HPM_HIP(h, hipStreamSynchronize(h->stream));
    return HPM_OK;
}

extern "C" int hpm_device_selftest(hpm_handle* h, int kind, const void* in, void* out, int n)
{
    if (!h || !in || !out || n < 0 || kind < 0 || kind > 4) {
        return fail(h, HPM_EINVAL, "bad selftest arguments");
    }
    if (n == 0) {
        return HPM_OK;
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const size_t in_bytes = (size_t)n * (kind == 4 ? 24 : 4);
    const size_t out_bytes = (size_t)n * (kind == 4 ? 16 : 4);
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
        return fail(h, HPM_EDEVICE, "selftest failed: %s", hipGetErrorString(e));
    }
    return HPM_OK;
}
