/*
 * hpm_prior_gpu.cpp — one scale of the planar prior built on the device from
 * the handle's last RunPatchMatch (main.cpp:84-162 at 0.25x, 238-314 at
 * 0.5x, 387-454 at 1x; HPM.cpp:759-1123): the scale maps, the support-point
 * blocks, the rasterised triangle mask, the triangle planes, the KNN
 * supplement and the prior maps run as kernels (hpm_prior.hip), as do the
 * gathering of the support points, their Delaunay triangulation (the star
 * walk of hpm_dt.h, which the host's hpm_delaunay also runs) and the KNN
 * search grid; only the point and triangle counts come back to size the
 * launches. The result equals the host construction of hpm_pipeline.cpp bit
 * for bit.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "hpm_internal.h"
#include "hpm_prior.h"

extern "C" {
int hpm_launch_prior_maps(const float* hyp4, const float* cost, int W, int H, float* depth_s, float* cost_s, int ws,
                          int hs, void* stream);
int hpm_launch_support_blocks(const float* cost, int w, int h, int32_t* cand, void* stream);
int hpm_launch_raster(const int32_t* pts, const int32_t* tris, const int32_t* tri_idx, int ntris, uint32_t* mask,
                      int width, int height, void* stream);
int hpm_launch_tri_planes(const int32_t* pts, const int32_t* tris, int ntris, const float* depth, int width,
                          const hpm_camera* ref, float factor, float* planes, void* stream);
int hpm_launch_knn(const int32_t* pts, int npoints, const int32_t* cell_start, const int32_t* cell_pts, int cell,
                   int gw, int gh, const int32_t* order, int norder, int k, const uint32_t* raster, const float* depth,
                   int width, int height, const hpm_camera* ref, float factor, float depth_min, float depth_max,
                   uint8_t* knn_flag, float* knn_planes, void* stream);
int hpm_launch_support_delaunay(const int32_t* cand, int ws, int hs, int cell, int gw, int gh, int32_t* pts,
                                int32_t* tris, int32_t* tri_idx, int32_t* cell_start, int32_t* cell_pts,
                                int32_t* ws_i32, int32_t* nums, void* stream);
int hpm_launch_prior_assemble(const uint32_t* raster, const float* tri_planes, const uint8_t* knn_flag,
                              const float* knn_planes, int width, int height, const hpm_camera* ref, float factor,
                              float depth_min, float depth_max, float* src_nd, float* planes, uint32_t* mask,
                              void* stream);
}

namespace {

/* the triple order of HPM.cpp:982-1084 (as hpm_knn_supplement) */
std::vector<int32_t> triple_order(int k)
{
    std::vector<int32_t> order;
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b)
            for (int c = b + 1; c < 5; ++c) {
                order.push_back(a);
                order.push_back(b);
                order.push_back(c);
            }
    for (int m = 5; m < k; ++m)
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b) {
                order.push_back(a);
                order.push_back(b);
                order.push_back(m);
            }
    return order;
}

struct Carve {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t n)
    {
        T* p = (T*)(base + off);
        off += (n * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

} // namespace

extern "C" int hpm_build_planar_prior(hpm_handle* h, float factor, int knn_k, double* ms_host)
{
    using clk = std::chrono::steady_clock;
    if (!h) {
        return HPM_EINVAL;
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "device initialisation must come first");
    }
    if (!(factor == 1.0f || factor == 0.5f || factor == 0.25f) || knn_k < 3 || knn_k > 32) {
        return hpm_fail(h, HPM_EINVAL, "planar prior: factor 0.25, 0.5 or 1 and knn_k 3..32 (got %g, %d)",
                        (double)factor, knn_k);
    }
    HPM_HIP(h, hipSetDevice(h->device));
    const hpm_camera ref = h->cams[0];
    const int W = ref.width, H = ref.height;
    const float dmin = h->params.depth_min, dmax = h->params.depth_max;
    const int ws = factor == 1.0f ? W : (int)lround(W * (double)factor);
    const int hs = factor == 1.0f ? H : (int)lround(H * (double)factor);
    const size_t ns = (size_t)ws * hs;
    const int nbx = (ws + 4) / 5, nby = (hs + 4) / 5, nb = nbx * nby;
    const int cell = 16;
    const int gw = (ws + cell - 1) / cell, gh = (hs + cell - 1) / cell;
    const std::vector<int32_t> order = triple_order(knn_k);
    const int norder = (int)(order.size() / 3);
    /* the handle's device state is used by the kernels below */
    if (!h->d_prior) {
        HPM_HIP(h, hipMalloc((void**)&h->d_prior, (size_t)W * H * 16));
        HPM_HIP(h, hipMalloc((void**)&h->d_mask, (size_t)W * H * 4));
    }
    /* workspace: a triangulation of n points has fewer than 2n triangles */
    const size_t maxp = (size_t)nb, maxt = 2 * (size_t)nb;
    auto layout = [&](Carve& c, float** depth_s, float** cost_s, int32_t** cand, int32_t** pts, int32_t** tris,
                      int32_t** tidx, uint32_t** raster, float** tplanes, int32_t** cstart, int32_t** cpts,
                      int32_t** ord, uint8_t** kflag, float** kplanes, float** src_nd, int32_t** scratch,
                      int32_t** nums) {
        *depth_s = c.take<float>(ns);
        *cost_s = c.take<float>(ns);
        *cand = c.take<int32_t>((size_t)nb);
        *pts = c.take<int32_t>(2 * maxp);
        *tris = c.take<int32_t>(3 * maxt);
        *tidx = c.take<int32_t>(maxt);
        *raster = c.take<uint32_t>(ns);
        *tplanes = c.take<float>(4 * maxt);
        *cstart = c.take<int32_t>((size_t)gw * gh + 1);
        *cpts = c.take<int32_t>(maxp);
        *ord = c.take<int32_t>(order.size());
        *kflag = c.take<uint8_t>(ns);
        *kplanes = c.take<float>(4 * ns);
        *src_nd = c.take<float>(factor == 1.0f ? 0 : 4 * ns);
        *scratch = c.take<int32_t>(21 * (size_t)nb + 10 * (size_t)ws + 16);
        *nums = c.take<int32_t>(4);
    };
    float *depth_s, *cost_s, *tplanes, *kplanes, *src_nd;
    int32_t *cand, *pts, *tris, *tidx, *cstart, *cpts, *ord, *scratch, *nums;
    uint32_t* raster;
    uint8_t* kflag;
    Carve sizing{nullptr};
    layout(sizing, &depth_s, &cost_s, &cand, &pts, &tris, &tidx, &raster, &tplanes, &cstart, &cpts, &ord, &kflag,
           &kplanes, &src_nd, &scratch, &nums);
    void* ws_ptr = nullptr;
    int rc = hpm_workspace(h, sizing.off, &ws_ptr);
    if (rc != HPM_OK) {
        return rc;
    }
    Carve carve{(char*)ws_ptr};
    layout(carve, &depth_s, &cost_s, &cand, &pts, &tris, &tidx, &raster, &tplanes, &cstart, &cpts, &ord, &kflag,
           &kplanes, &src_nd, &scratch, &nums);
    hipStream_t s = h->stream;
    HPM_HIP(h, (hipError_t)hpm_launch_prior_maps(h->d_hyp, h->d_cost, W, H, depth_s, cost_s, ws, hs, s));
    HPM_HIP(h, (hipError_t)hpm_launch_support_blocks(cost_s, ws, hs, cand, s));
    HPM_HIP(h, (hipError_t)hpm_launch_support_delaunay(cand, ws, hs, cell, gw, gh, pts, tris, tidx, cstart, cpts,
                                                        scratch, nums, s));
    /* the counts size the per-triangle and per-point launches */
    int32_t h_nums[3] = {0, 0, 0};
    auto t0 = clk::now();
    HPM_HIP(h, hipMemcpyAsync(h_nums, nums, sizeof(h_nums), hipMemcpyDeviceToHost, s));
    HPM_HIP(h, hipStreamSynchronize(s));
    if (ms_host) {
        *ms_host += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    if (h_nums[2] != 0 || h_nums[1] > (int)maxt) {
        return hpm_fail(h, HPM_EDEVICE, "planar prior: the Delaunay star walk failed (flags %d, %d triangles)",
                        h_nums[2], h_nums[1]);
    }
    const int npts = h_nums[0];
    const int ntris = h_nums[1];
    HPM_HIP(h, hipMemcpyAsync(ord, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
    HPM_HIP(h, (hipError_t)hpm_launch_raster(pts, tris, tidx, ntris, raster, ws, hs, s));
    HPM_HIP(h, (hipError_t)hpm_launch_tri_planes(pts, tris, ntris, depth_s, ws, &ref, factor, tplanes, s));
    HPM_HIP(h, (hipError_t)hpm_launch_knn(pts, npts, cstart, cpts, cell, gw, gh, ord, norder, knn_k, raster, depth_s,
                                          ws, hs, &ref, factor, dmin, dmax, kflag, kplanes, s));
    if (factor == 1.0f) {
        HPM_HIP(h, (hipError_t)hpm_launch_prior_assemble(raster, tplanes, kflag, kplanes, ws, hs, &ref, factor, dmin,
                                                         dmax, nullptr, h->d_prior, h->d_mask, s));
    } else {
        HPM_HIP(h, (hipError_t)hpm_launch_prior_assemble(raster, tplanes, kflag, kplanes, ws, hs, &ref, factor, dmin,
                                                         dmax, src_nd, nullptr, nullptr, s));
        const int scale_i = std::max(H / hs, W / ws);
        if (scale_i == 1) {
            /* JointBilateralUpsampling returns without output at integer
             * scale 1 (HPM.cpp:1219-1222): zero prior */
            HPM_HIP(h, hipMemsetAsync(h->d_prior, 0, (size_t)W * H * 16, s));
            HPM_HIP(h, hipMemsetAsync(h->d_mask, 0, (size_t)W * H * 4, s));
        } else {
            HPM_HIP(h, (hipError_t)hpm_launch_jbu_prior(h->pool->d_images + h->pool->off[(size_t)h->views[0]], W, H,
                                                        src_nd, ws, hs, scale_i, &ref, dmin, dmax, h->d_prior,
                                                        h->d_mask, s));
        }
    }
    HPM_HIP(h, hipStreamSynchronize(s));
    h->have_prior = true;
    h->params.planar_prior = 1;
    return HPM_OK;
}
