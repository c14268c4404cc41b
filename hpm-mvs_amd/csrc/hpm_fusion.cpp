/*
 * hpm_fusion.cpp — RunFusion (HPM-MVS/main.cpp:502-633) over a pool and the
 * binary PLY writer (HPM.cpp:403-456).
 *
 * The consistency tests of every (reference pixel, source) pair are pure
 * functions of the depth and normal maps and run on the device
 * (hpm_k_fusion_geometry). What follows them is sequential in the
 * reference: a kept point masks source pixels (the last consistent pixel of
 * each source so far in this view, used_list, which the reference resets
 * per view, not per pixel), and later pixels skip masked sources. That scan
 * replays the records on the host in the reference's order, so the point
 * list is the reference's for the same maps. It is integer and compare work
 * over one record per (pixel, source), with the pixels no source confirms
 * skipped by a per-pixel bitmask.
 */
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hpm_fusion.h"
#include "hpm_internal.h"
#include "hpm_io.h"
#include "hpm_pipeline.h"

namespace {

struct P3 {
    float x, y, z;
};

/* Get3DPointonWorld (HPM.cpp:209-233), the device kernel's arithmetic */
P3 world_point(int x, int y, float depth, const hpm_camera& c)
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

int fail(char* err, int err_len, const char* fmt, ...)
{
    if (err && err_len > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, (size_t)err_len, fmt, ap);
        va_end(ap);
    }
    return HPM_EINVAL;
}

uint8_t grey_byte(float v)
{
    const int i = (int)v;
    return (uint8_t)(i < 0 ? 0 : i > 255 ? 255 : i);
}

/* RunFusion over the pool: HPM-MVS's rule, or with acmm ACMM+NESP's
 * (ACMM+NESP/windows/main.cpp:164-300: the point, normal and colour are the
 * means of the reference pixel's and its consistent sources', the normal
 * through cv::Vec3f /= float, a product with the float reciprocal; a point
 * needs 2 consistent sources) */
/* colour channel k of pixel p of view v: the given RGB image, else the grey
 * image in every channel */
inline float colour_of(const hpm_pool* pool, const uint8_t* const* colors, int v, size_t p, int k)
{
    return colors && colors[v] ? (float)colors[v][3 * p + (size_t)k] : (float)grey_byte(pool->images[(size_t)v][p]);
}

int fusion_run(bool acmm, hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
               int geom_consistency, const uint8_t* const* colors, hpm_point** points, int64_t* num_points, char* err,
               int err_len)
{
    if (!pool || !num_src || !src_ids || !points || !num_points || num_views < 1 || num_views > pool->n ||
        max_src < 1) {
        return fail(err, err_len, "fusion: bad arguments");
    }
    *points = nullptr;
    *num_points = 0;
    size_t max_px = 0;
    int max_nn = 0;
    for (int i = 0; i < num_views; ++i) {
        if (num_src[i] < 0 || num_src[i] > 32 || num_src[i] > max_src) {
            return fail(err, err_len, "fusion: view %d has %d sources (at most 32 and max_src)", i, num_src[i]);
        }
        for (int j = 0; j < num_src[i]; ++j) {
            const int s = src_ids[(size_t)i * max_src + j];
            if (s < 0 || s >= num_views || s == i) {
                return fail(err, err_len, "fusion: view %d source %d is %d", i, j, s);
            }
        }
        if (!pool->have_state[(size_t)i]) {
            return fail(err, err_len, "fusion: view %d has no depth map in the pool", i);
        }
        max_px = std::max(max_px, (size_t)pool->cams[(size_t)i].width * pool->cams[(size_t)i].height);
        max_nn = std::max(max_nn, num_src[i]);
    }
    if (hipSetDevice(pool->device) != hipSuccess) {
        return fail(err, err_len, "fusion: hipSetDevice failed");
    }
    const size_t nrec = max_px * (size_t)std::max(max_nn, 1);
    hpm_camera* d_cams = nullptr;
    int64_t* d_off = nullptr;
    int32_t* d_pix = nullptr;
    float* d_w = nullptr;
    uint32_t* d_any = nullptr;
    int32_t* h_pix = nullptr;
    float* h_w = nullptr;
    uint32_t* h_any = nullptr;
    float* h_depth = nullptr;
    float* h_hyp = nullptr;
    hipStream_t s = pool->stream;
    bool ok = hipMalloc((void**)&d_cams, sizeof(hpm_camera) * (size_t)pool->n) == hipSuccess &&
              hipMalloc((void**)&d_off, sizeof(int64_t) * (size_t)pool->n) == hipSuccess &&
              hipMalloc((void**)&d_pix, nrec * 4) == hipSuccess && hipMalloc((void**)&d_w, nrec * 4) == hipSuccess &&
              hipMalloc((void**)&d_any, max_px * 4) == hipSuccess &&
              hipHostMalloc((void**)&h_pix, nrec * 4, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&h_w, nrec * 4, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&h_any, max_px * 4, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&h_depth, max_px * 4, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&h_hyp, max_px * 16, hipHostMallocDefault) == hipSuccess &&
              hipMemcpyAsync(d_cams, pool->cams.data(), sizeof(hpm_camera) * (size_t)pool->n,
                             hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_off, pool->off.data(), sizeof(int64_t) * (size_t)pool->n, hipMemcpyHostToDevice,
                             s) == hipSuccess;
    const float* depth_pool = geom_consistency ? pool->d_depth_geom : pool->d_depth_plain;
    std::vector<std::vector<uint8_t>> masks((size_t)num_views);
    for (int i = 0; i < num_views; ++i) {
        masks[(size_t)i].assign((size_t)pool->cams[(size_t)i].width * pool->cams[(size_t)i].height, 0);
    }
    /* ACMM: every view's depth and normals on the host, for the sums */
    std::vector<float> all_depth, all_hyp;
    if (ok && acmm) {
        all_depth.resize((size_t)pool->total);
        all_hyp.resize((size_t)pool->total * 4);
        ok = hipMemcpyAsync(all_depth.data(), depth_pool, all_depth.size() * 4, hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipMemcpyAsync(all_hyp.data(), pool->d_state_hyp, all_hyp.size() * 4, hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    std::vector<hpm_point> out;
    for (int i = 0; ok && i < num_views; ++i) {
        const hpm_camera& rc = pool->cams[(size_t)i];
        const int cols = rc.width, rows = rc.height;
        const size_t npx = (size_t)cols * rows;
        const int nn = num_src[i];
        const int* srcs = src_ids + (size_t)i * max_src;
        hpm_fusion_args a;
        memset(&a, 0, sizeof(a));
        a.cams = d_cams;
        a.off = d_off;
        a.depth = depth_pool;
        a.hyp = pool->d_state_hyp;
        a.ref = i;
        a.nn = nn;
        for (int j = 0; j < nn; ++j) {
            a.src[j] = srcs[j];
        }
        a.rec_pix = d_pix;
        a.rec_w = d_w;
        a.any = d_any;
        ok = hpm_launch_fusion_geometry(&a, cols, rows, s) == 0 &&
             hipMemcpyAsync(h_any, d_any, npx * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(h_pix, d_pix, npx * (size_t)nn * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(h_w, d_w, npx * (size_t)nn * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(h_depth, depth_pool + pool->off[(size_t)i], npx * 4, hipMemcpyDeviceToHost, s) ==
                 hipSuccess &&
             hipMemcpyAsync(h_hyp, pool->d_state_hyp + 4 * pool->off[(size_t)i], npx * 16, hipMemcpyDeviceToHost,
                            s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        if (!ok) {
            break;
        }
        /* the reference's scan (main.cpp:553-629) over the device records */
        int ux[32], uy[32];
        for (int j = 0; j < nn; ++j) {
            ux[j] = -1;
            uy[j] = -1;
        }
        std::vector<uint8_t>& own = masks[(size_t)i];
        for (size_t p = 0; p < npx; ++p) {
            const uint32_t any = h_any[p];
            if (any == 0u || own[p] == 1) {
                continue; /* no source confirms the pixel (or depth <= 0), or it is masked */
            }
            int num_consistent = 0;
            float dyn = 0.0f;
            const int c = (int)(p % (size_t)cols), r = (int)(p / (size_t)cols);
            P3 sx{0.0f, 0.0f, 0.0f};
            float sn[3] = {h_hyp[4 * p], h_hyp[4 * p + 1], h_hyp[4 * p + 2]};
            float sc[3] = {colour_of(pool, colors, i, p, 0), colour_of(pool, colors, i, p, 1),
                           colour_of(pool, colors, i, p, 2)};
            if (acmm) {
                sx = world_point(c, r, h_depth[p], rc);
            }
            for (int j = 0; j < nn; ++j) {
                if (((any >> j) & 1u) == 0u) {
                    continue;
                }
                const int32_t q = h_pix[p * nn + j];
                const int sv = srcs[j];
                if (masks[(size_t)sv][(size_t)q] == 1) {
                    continue;
                }
                const int sw = pool->cams[(size_t)sv].width;
                ux[j] = q % sw;
                uy[j] = q / sw;
                if (acmm) {
                    /* the source's point as the consistency test built it */
                    const size_t g = (size_t)pool->off[(size_t)sv] + (size_t)q;
                    const P3 Y = world_point(ux[j], uy[j], all_depth[g], pool->cams[(size_t)sv]);
                    sx.x += Y.x;
                    sx.y += Y.y;
                    sx.z += Y.z;
                    sn[0] = sn[0] + all_hyp[4 * g];
                    sn[1] = sn[1] + all_hyp[4 * g + 1];
                    sn[2] = sn[2] + all_hyp[4 * g + 2];
                    for (int k = 0; k < 3; ++k) {
                        sc[k] += colour_of(pool, colors, sv, (size_t)q, k);
                    }
                }
                dyn += h_w[p * nn + j];
                ++num_consistent;
            }
            const bool keep =
                acmm ? num_consistent >= 2 : (num_consistent >= 1 && (double)dyn > 0.3 * (double)num_consistent);
            if (keep) {
                hpm_point pt;
                if (acmm) {
                    const float k = num_consistent + 1.0f;
                    const float ik = 1.0f / k;
                    pt.x = sx.x / k;
                    pt.y = sx.y / k;
                    pt.z = sx.z / k;
                    pt.nx = sn[0] * ik;
                    pt.ny = sn[1] * ik;
                    pt.nz = sn[2] * ik;
                    pt.r = (uint8_t)(int)(sc[0] / k);
                    pt.g = (uint8_t)(int)(sc[1] / k);
                    pt.b = (uint8_t)(int)(sc[2] / k);
                } else {
                    const P3 X = world_point(c, r, h_depth[p], rc);
                    pt.x = X.x;
                    pt.y = X.y;
                    pt.z = X.z;
                    pt.nx = sn[0];
                    pt.ny = sn[1];
                    pt.nz = sn[2];
                    pt.r = (uint8_t)(int)sc[0];
                    pt.g = (uint8_t)(int)sc[1];
                    pt.b = (uint8_t)(int)sc[2];
                }
                pt.pad_ = 0;
                out.push_back(pt);
                for (int j = 0; j < nn; ++j) {
                    if (ux[j] == -1) {
                        continue;
                    }
                    const int sv = srcs[j];
                    masks[(size_t)sv][(size_t)uy[j] * pool->cams[(size_t)sv].width + ux[j]] = 1;
                }
            }
        }
    }
    (void)hipFree(d_cams);
    (void)hipFree(d_off);
    (void)hipFree(d_pix);
    (void)hipFree(d_w);
    (void)hipFree(d_any);
    (void)hipHostFree(h_pix);
    (void)hipHostFree(h_w);
    (void)hipHostFree(h_any);
    (void)hipHostFree(h_depth);
    (void)hipHostFree(h_hyp);
    if (!ok) {
        return fail(err, err_len, "fusion: device error");
    }
    hpm_point* arr = (hpm_point*)malloc(sizeof(hpm_point) * std::max(out.size(), (size_t)1));
    if (!arr) {
        return fail(err, err_len, "fusion: out of host memory");
    }
    if (!out.empty()) {
        memcpy(arr, out.data(), sizeof(hpm_point) * out.size());
    }
    *points = arr;
    *num_points = (int64_t)out.size();
    return HPM_OK;
}

} // namespace

extern "C" int hpm_fusion_run(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
                              int geom_consistency, hpm_point** points, int64_t* num_points, char* err, int err_len)
{
    return fusion_run(false, pool, num_views, num_src, src_ids, max_src, geom_consistency, nullptr, points,
                      num_points, err, err_len);
}

extern "C" int hpm_fusion_run_acmm(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids,
                                   int max_src, int geom_consistency, hpm_point** points, int64_t* num_points,
                                   char* err, int err_len)
{
    return fusion_run(true, pool, num_views, num_src, src_ids, max_src, geom_consistency, nullptr, points,
                      num_points, err, err_len);
}

extern "C" int hpm_fusion_run_rgb(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
                                  int geom_consistency, const uint8_t* const* colors, int acmm, hpm_point** points,
                                  int64_t* num_points, char* err, int err_len)
{
    return fusion_run(acmm != 0, pool, num_views, num_src, src_ids, max_src, geom_consistency, colors, points,
                      num_points, err, err_len);
}

extern "C" void hpm_fusion_free(hpm_point* points)
{
    free(points);
}

extern "C" int hpm_write_ply(const char* path, const hpm_point* points, int64_t num_points)
{
    if (!path || (num_points > 0 && !points) || num_points < 0) {
        return -1;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        return -1;
    }
    fprintf(f,
            "ply\nformat binary_little_endian 1.0\nelement vertex %lld\nproperty float x\nproperty float y\n"
            "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty uchar red\n"
            "property uchar green\nproperty uchar blue\nend_header\n",
            (long long)num_points);
    bool ok = true;
    std::vector<unsigned char> buf;
    const int64_t chunk = 65536;
    for (int64_t i0 = 0; ok && i0 < num_points; i0 += chunk) {
        const int64_t n = std::min(chunk, num_points - i0);
        buf.resize((size_t)n * 27);
        for (int64_t k = 0; k < n; ++k) {
            const hpm_point& p = points[i0 + k];
            float v[6] = {p.x, p.y, p.z, p.nx, p.ny, p.nz};
            if (!(v[0] < FLT_MAX && v[0] > -FLT_MAX) || !(v[1] < FLT_MAX && v[1] > -FLT_MAX) ||
                !(v[2] < FLT_MAX && v[2] >= -FLT_MAX)) {
                v[0] = v[1] = v[2] = 0.0f;
            }
            unsigned char* d = buf.data() + (size_t)k * 27;
            memcpy(d, v, 24);
            d[24] = p.r;
            d[25] = p.g;
            d[26] = p.b;
        }
        ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

namespace {

/* RunFusion as main calls it, over the maps of <dense_folder>/<sub> */
int fusion_folder(bool acmm, const char* dense_folder, int geom_consistency, int64_t* num_points, char* err,
                  int err_len)
{
    const char* sub = acmm ? "ACMM_NESP" : "HPM";
    if (!dense_folder) {
        return fail(err, err_len, "fusion: null folder");
    }
    const std::string base(dense_folder);
    int np = 0;
    if (hpm_read_pair((base + "/pair.txt").c_str(), &np, nullptr, nullptr, nullptr, 0) != 0 || np < 1) {
        return fail(err, err_len, "fusion: cannot read %s/pair.txt", dense_folder);
    }
    std::vector<int> refs((size_t)np), ns((size_t)np), srcs((size_t)np * HPM_MAX_SOURCES);
    hpm_read_pair((base + "/pair.txt").c_str(), &np, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES);
    for (int i = 0; i < np; ++i) {
        if (refs[(size_t)i] != i) { /* RunFusion indexes the maps by problem (main.cpp:517-547) */
            return fail(err, err_len, "fusion: pair.txt must list views 0..%d in order", np - 1);
        }
    }
    std::vector<hpm_camera> cams((size_t)np);
    std::vector<std::vector<float>> imgs((size_t)np), depth((size_t)np), hyp((size_t)np);
    std::vector<std::vector<uint8_t>> rgb((size_t)np);
    char path[4096];
    for (int i = 0; i < np; ++i) {
        int w = 0, h = 0;
        snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder, i);
        if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, i);
            if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
                return fail(err, err_len, "fusion: cannot read image %d (JPEG or 8-bit PGM expected)", i);
            }
        }
        imgs[(size_t)i].resize((size_t)w * h);
        hpm_read_image_grey(path, imgs[(size_t)i].data(), &w, &h);
        /* the colours, cv::imread(IMREAD_COLOR) (main.cpp:523) */
        rgb[(size_t)i].resize((size_t)w * h * 3);
        int cw = w, ch = h;
        if (hpm_read_image_rgb(path, rgb[(size_t)i].data(), &cw, &ch) != 0 || cw != w || ch != h) {
            return fail(err, err_len, "fusion: cannot read the colours of image %d", i);
        }
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, i);
        if (hpm_read_camera(path, &cams[(size_t)i]) != 0) {
            return fail(err, err_len, "fusion: cannot read %s", path);
        }
        cams[(size_t)i].width = w;
        cams[(size_t)i].height = h;
    }
    std::vector<const float*> ip;
    for (auto& v : imgs) {
        ip.push_back(v.data());
    }
    hpm_params defaults;
    hpm_params_default(&defaults);
    hpm_pool* pool = hpm_pool_create(0, np, cams.data(), ip.data(), defaults.max_image_size);
    if (!pool) {
        return fail(err, err_len, "fusion: cannot create the device pool");
    }
    int rc = HPM_OK;
    std::vector<const uint8_t*> cp((size_t)np);
    for (int i = 0; i < np; ++i) {
        /* a view the pool downscaled (max_image_size): its colours resized
         * as RescaleImageAndCamera does (main.cpp:541-542), with the
         * library's bilinear resize rounded to bytes */
        const hpm_camera& pc = pool->cams[(size_t)i];
        if (pc.width != cams[(size_t)i].width || pc.height != cams[(size_t)i].height) {
            const int w = cams[(size_t)i].width, h = cams[(size_t)i].height;
            std::vector<float> f((size_t)w * h * 3), g((size_t)pc.width * pc.height * 3);
            for (size_t k = 0; k < f.size(); ++k) f[k] = (float)rgb[(size_t)i][k];
            hpm_resize_bilinear_c(f.data(), w, h, 3, g.data(), pc.width, pc.height);
            rgb[(size_t)i].resize(g.size());
            for (size_t k = 0; k < g.size(); ++k) {
                const long v = lrintf(g[k]);
                rgb[(size_t)i][k] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
            }
        }
        cp[(size_t)i] = rgb[(size_t)i].data();
    }
    for (int i = 0; i < np && rc == HPM_OK; ++i) {
        const hpm_camera& pc = pool->cams[(size_t)i];
        const size_t n = (size_t)pc.width * pc.height;
        std::vector<float> d(n), nrm(n * 3), h4(n * 4);
        int w = 0, h = 0, ch = 0;
        snprintf(path, sizeof(path), "%s/%s/2333_%08d/%s", dense_folder, sub, i,
                 geom_consistency ? "depths_geom.dmb" : "depths.dmb");
        bool ok = hpm_read_dmb(path, nullptr, &w, &h, &ch) == 0 && w == pc.width && h == pc.height && ch == 1 &&
                  hpm_read_dmb(path, d.data(), &w, &h, &ch) == 0;
        snprintf(path, sizeof(path), "%s/%s/2333_%08d/normals.dmb", dense_folder, sub, i);
        ok = ok && hpm_read_dmb(path, nullptr, &w, &h, &ch) == 0 && w == pc.width && h == pc.height && ch == 3 &&
             hpm_read_dmb(path, nrm.data(), &w, &h, &ch) == 0;
        if (!ok) {
            rc = fail(err, err_len, "fusion: missing or mis-sized maps of view %d (%s)", i, path);
            break;
        }
        for (size_t k = 0; k < n; ++k) {
            h4[4 * k] = nrm[3 * k];
            h4[4 * k + 1] = nrm[3 * k + 1];
            h4[4 * k + 2] = nrm[3 * k + 2];
            h4[4 * k + 3] = d[k];
        }
        if (hpm_pool_set_depth(pool, i, geom_consistency ? 1 : 0, d.data()) != HPM_OK ||
            hpm_pool_set_state(pool, i, h4.data(), nullptr) != HPM_OK) {
            rc = fail(err, err_len, "fusion: device error");
        }
    }
    hpm_point* pts = nullptr;
    int64_t n = 0;
    if (rc == HPM_OK) {
        rc = hpm_fusion_run_rgb(pool, np, ns.data(), srcs.data(), HPM_MAX_SOURCES, geom_consistency, cp.data(),
                                acmm ? 1 : 0, &pts, &n, err, err_len);
    }
    hpm_pool_destroy(pool);
    if (rc == HPM_OK) {
        const std::string ply = base + (acmm ? "/ACMM_NESP/ACMM_NESP_model.ply" : "/HPM/HPM_model.ply");
        if (hpm_write_ply(ply.c_str(), pts, n) != 0) {
            rc = fail(err, err_len, "fusion: cannot write %s", ply.c_str());
        }
    }
    hpm_fusion_free(pts);
    if (rc == HPM_OK && num_points) {
        *num_points = n;
    }
    return rc;
}

} // namespace

extern "C" int hpm_fusion_run_folder(const char* dense_folder, int geom_consistency, int64_t* num_points, char* err,
                                     int err_len)
{
    return fusion_folder(false, dense_folder, geom_consistency, num_points, err, err_len);
}

extern "C" int hpm_fusion_run_folder_acmm(const char* dense_folder, int geom_consistency, int64_t* num_points,
                                          char* err, int err_len)
{
    return fusion_folder(true, dense_folder, geom_consistency, num_points, err, err_len);
}
