/*
 * hpm_acmm.cpp — the ACMM+NESP coarse-to-fine schedule
 * (ACMM+NESP/windows/main.cpp:35-378): one device pool per level, each
 * level run by the pass schedule of hpm_pipeline.cpp with the ACMM kernel
 * variant (no planar prior, geometric passes of 2 iterations), the levels
 * after the first starting from the previous level's final maps through the
 * hierarchy initialisation.
 */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "hpm_acmm.h"
#include "hpm_fusion.h"
#include "hpm_internal.h"
#include "hpm_io.h"

namespace {

int fail(char* err, int err_len, int code, const char* fmt, ...)
{
    if (err && err_len > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, (size_t)err_len, fmt, ap);
        va_end(ap);
    }
    return code;
}

/* one view scaled to a level: ACMM::InuputInitialization's rescale
 * (ACMM.cpp:497-523); unchanged when it fits cur */
void scale_view(const hpm_camera& cam, const float* img, int cur, hpm_camera* out_cam, std::vector<float>* out)
{
    const int W = cam.width, H = cam.height;
    *out_cam = cam;
    if (W <= cur && H <= cur) {
        out->assign(img, img + (size_t)W * H);
        return;
    }
    const float fx = (float)cur / W;
    const float fy = (float)cur / H;
    const float f = std::min(fx, fy);
    const int nw = (int)std::round(W * f);
    const int nh = (int)std::round(H * f);
    out->resize((size_t)nw * nh);
    hpm_resize_bilinear(img, W, H, out->data(), nw, nh);
    const float sx = nw / (float)W, sy = nh / (float)H;
    out_cam->K[0] *= sx;
    out_cam->K[2] *= sx;
    out_cam->K[4] *= sy;
    out_cam->K[5] *= sy;
    out_cam->width = nw;
    out_cam->height = nh;
}

void add_stats(hpm_pipeline_stats* to, const hpm_pipeline_stats& s)
{
    to->ms_phase_a += s.ms_phase_a;
    to->ms_phase_b += s.ms_phase_b;
    to->ms_device += s.ms_device;
    to->n_eval += s.n_eval;
    to->n_eval_propagate += s.n_eval_propagate;
    to->ms_propagate += s.ms_propagate;
    to->runs += s.runs;
    to->propagate_launches += s.propagate_launches;
}

} // namespace

extern "C" void hpm_acmm_default(hpm_acmm_options* opt)
{
    memset(opt, 0, sizeof(*opt));
    hpm_params_default(&opt->params);
    opt->params.acmm = 1;
    opt->scale_bound = 1000;
    opt->geom_iterations = 2;
    opt->workers = 12;
}

extern "C" int hpm_acmm_run(int device, int num_views, const hpm_camera* cams, const float* const* images,
                            int num_problems, const int* ref_ids, const int* num_src, const int* src_ids, int max_src,
                            const hpm_acmm_options* opt, hpm_pool** out_pool, hpm_pipeline_stats* stats, char* err,
                            int err_len)
{
    hpm_acmm_options defaults;
    hpm_acmm_default(&defaults);
    const hpm_acmm_options* o = opt ? opt : &defaults;
    if (!out_pool || num_views < 1 || !cams || !images || num_problems < 1 || !ref_ids || !num_src || !src_ids ||
        max_src < 1) {
        return fail(err, err_len, HPM_EINVAL, "bad ACMM arguments");
    }
    *out_pool = nullptr;
    if (!o->params.acmm || o->params.planar_prior || o->scale_bound < 1 || o->geom_iterations < 0) {
        return fail(err, err_len, HPM_EINVAL, "ACMM options need params.acmm, no planar prior, scale_bound >= 1");
    }
    /* the problem of each view: its level sizes (ACMM.cpp:497-500 indexes
     * the problems by source image id) */
    std::vector<int> prob_of((size_t)num_views, -1);
    for (int i = 0; i < num_problems; ++i) {
        const int r = ref_ids[i];
        if (r < 0 || r >= num_views || prob_of[(size_t)r] >= 0) {
            return fail(err, err_len, HPM_EINVAL, "problem %d: bad or repeated reference view %d", i, r);
        }
        prob_of[(size_t)r] = i;
    }
    for (int i = 0; i < num_problems; ++i) {
        for (int j = 0; j < num_src[i]; ++j) {
            const int s = src_ids[(size_t)i * max_src + j];
            if (s < 0 || s >= num_views || prob_of[(size_t)s] < 0) {
                return fail(err, err_len, HPM_EINVAL, "source view %d of problem %d is no problem's reference", s, i);
            }
        }
    }
    /* ComputeMultiScaleSettings (main.cpp:35-72) */
    std::vector<int> max_size((size_t)num_problems), left((size_t)num_problems), cur((size_t)num_problems);
    int levels = -1;
    for (int i = 0; i < num_problems; ++i) {
        const hpm_camera& c = cams[ref_ids[i]];
        int m = std::min(std::max(c.width, c.height), o->params.max_image_size);
        max_size[(size_t)i] = m;
        int k = 0;
        while (m > o->scale_bound) {
            m /= 2;
            ++k;
        }
        left[(size_t)i] = k;
        cur[(size_t)i] = max_size[(size_t)i];
        levels = std::max(levels, k);
    }
    hpm_pipeline_stats total;
    memset(&total, 0, sizeof(total));
    hpm_pool* prev = nullptr;
    int rc = HPM_OK;
    for (int level = levels, step = 0; level >= 0 && rc == HPM_OK; --level, ++step) {
        /* main.cpp:319-324: cur_image_size = max_image_size / 2^num_downscale */
        for (int i = 0; i < num_problems; ++i) {
            if (left[(size_t)i] >= 0) {
                cur[(size_t)i] = (int)(max_size[(size_t)i] / std::pow(2.0, left[(size_t)i]));
                left[(size_t)i] -= 1;
            }
        }
        std::vector<hpm_camera> lc((size_t)num_views);
        std::vector<std::vector<float>> li((size_t)num_views);
        std::vector<const float*> lp((size_t)num_views);
        for (int v = 0; v < num_views; ++v) {
            const int p = prob_of[(size_t)v];
            const int size = p >= 0 ? cur[(size_t)p] : std::max(cams[v].width, cams[v].height);
            scale_view(cams[v], images[v], size, &lc[(size_t)v], &li[(size_t)v]);
            lp[(size_t)v] = li[(size_t)v].data();
        }
        if (prev) {
            for (int v = 0; v < num_views; ++v) {
                if (prob_of[(size_t)v] >= 0 && lc[(size_t)v].width == prev->cams[(size_t)v].width &&
                    lc[(size_t)v].height == prev->cams[(size_t)v].height) {
                    /* the reference would reuse a stale depths.dmb here
                     * (RunJBU returns at scale 1, ACMM.cpp:809-812) */
                    rc = fail(err, err_len, HPM_EINVAL,
                              "view %d keeps its size %dx%d from one level to the next; the hierarchy needs every view "
                              "to change scale",
                              v, lc[(size_t)v].width, lc[(size_t)v].height);
                    break;
                }
            }
            if (rc != HPM_OK) {
                break;
            }
        }
        hpm_pool* pool = hpm_pool_create(device, num_views, lc.data(), lp.data(), 0);
        if (!pool) {
            rc = fail(err, err_len, HPM_EDEVICE, "cannot create the device pool of level %d", level);
            break;
        }
        hpm_pipeline_options po;
        hpm_pipeline_default(&po);
        po.params = o->params;
        po.params.hierarchy = 0;
        po.params.geom_iterations = o->geom_iterations;
        po.planar_prior = 0;
        po.geom_consistency = o->geom_iterations > 0;
        po.workers = o->workers;
        po.verbose = o->verbose;
        po.hierarchy_from = prev;
        po.stream_base = 16 * step; /* distinct Philox streams per level */
        if (o->verbose) {
            fprintf(stderr, "ACMM level %d: %s pass, %d geometric passes\n", level, prev ? "hierarchy" : "plain",
                    o->geom_iterations);
        }
        hpm_pipeline_stats st;
        rc = hpm_pipeline_run(pool, num_problems, ref_ids, num_src, src_ids, max_src, &po, &st, err, err_len);
        add_stats(&total, st);
        if (prev) {
            hpm_pool_destroy(prev);
        }
        prev = pool;
    }
    if (rc != HPM_OK) {
        if (prev) {
            hpm_pool_destroy(prev);
        }
        return rc;
    }
    if (stats) {
        *stats = total;
    }
    *out_pool = prev;
    return HPM_OK;
}

extern "C" int hpm_acmm_run_folder(const char* dense_folder, const hpm_acmm_options* opt, int fuse,
                                   int64_t* num_points, hpm_pipeline_stats* stats, char* err, int err_len)
{
    if (!dense_folder) {
        return fail(err, err_len, HPM_EINVAL, "null folder");
    }
    const std::string base(dense_folder);
    int np = 0;
    if (hpm_read_pair((base + "/pair.txt").c_str(), &np, nullptr, nullptr, nullptr, 0) != 0 || np < 1) {
        return fail(err, err_len, HPM_EIO, "cannot read %s/pair.txt", dense_folder);
    }
    std::vector<int> refs((size_t)np), ns((size_t)np), srcs((size_t)np * HPM_MAX_SOURCES);
    hpm_read_pair((base + "/pair.txt").c_str(), &np, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES);
    for (int i = 0; i < np; ++i) {
        if (refs[(size_t)i] != i) { /* the problems are indexed by image id (ACMM.cpp:497-500, main.cpp:180-230) */
            return fail(err, err_len, HPM_EINVAL, "pair.txt must list views 0..%d in order", np - 1);
        }
    }
    std::vector<hpm_camera> cams((size_t)np);
    std::vector<std::vector<float>> imgs((size_t)np);
    char path[4096];
    for (int id = 0; id < np; ++id) {
        int w = 0, hh = 0;
        snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder, id);
        if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, id);
            if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
                return fail(err, err_len, HPM_EIO, "cannot read image %d (JPEG or 8-bit PGM expected)", id);
            }
        }
        imgs[(size_t)id].resize((size_t)w * hh);
        hpm_read_image_grey(path, imgs[(size_t)id].data(), &w, &hh);
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, id);
        if (hpm_read_camera(path, &cams[(size_t)id]) != 0) {
            return fail(err, err_len, HPM_EIO, "cannot read %s", path);
        }
        cams[(size_t)id].width = w;
        cams[(size_t)id].height = hh;
    }
    std::vector<const float*> ip;
    for (auto& v : imgs) {
        ip.push_back(v.data());
    }
    hpm_pool* pool = nullptr;
    int rc = hpm_acmm_run(0, np, cams.data(), ip.data(), np, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES,
                          opt, &pool, stats, err, err_len);
    if (rc != HPM_OK) {
        return rc;
    }
    const std::string out = base + "/ACMM_NESP";
    mkdir(out.c_str(), 0755);
    const bool geom = opt == nullptr || opt->geom_iterations > 0;
    for (int v = 0; v < np && rc == HPM_OK; ++v) {
        const int W = pool->cams[(size_t)v].width, H = pool->cams[(size_t)v].height;
        const size_t n = (size_t)W * H;
        std::vector<float> hyp(n * 4), cost(n), nrm(n * 3), d_first(n), d_geom(n);
        if (hpm_pool_get_state(pool, v, hyp.data(), cost.data()) != HPM_OK ||
            hpm_pool_get_depth(pool, v, 0, d_first.data()) != HPM_OK ||
            hpm_pool_get_depth(pool, v, 1, d_geom.data()) != HPM_OK) {
            rc = fail(err, err_len, HPM_EDEVICE, "cannot read the maps of view %d", v);
            break;
        }
        for (size_t k = 0; k < n; ++k) {
            nrm[3 * k + 0] = hyp[4 * k + 0];
            nrm[3 * k + 1] = hyp[4 * k + 1];
            nrm[3 * k + 2] = hyp[4 * k + 2];
        }
        snprintf(path, sizeof(path), "%s/2333_%08d", out.c_str(), v);
        mkdir(path, 0755);
        const std::string d(path);
        if (hpm_write_dmb((d + "/depths.dmb").c_str(), d_first.data(), W, H, 1) != 0 ||
            hpm_write_dmb((d + "/normals.dmb").c_str(), nrm.data(), W, H, 3) != 0 ||
            hpm_write_dmb((d + "/costs.dmb").c_str(), cost.data(), W, H, 1) != 0 ||
            (geom && hpm_write_dmb((d + "/depths_geom.dmb").c_str(), d_geom.data(), W, H, 1) != 0)) {
            rc = fail(err, err_len, HPM_EIO, "cannot write the maps of view %d", v);
        }
    }
    if (rc == HPM_OK && fuse) {
        /* RunFusion on the last level's geometric maps (main.cpp:164-300,
         * 376-377) */
        hpm_point* pts = nullptr;
        int64_t npts = 0;
        rc = hpm_fusion_run_acmm(pool, np, ns.data(), srcs.data(), HPM_MAX_SOURCES, geom ? 1 : 0, &pts, &npts, err,
                                 err_len);
        if (rc == HPM_OK) {
            if (hpm_write_ply((out + "/ACMM_NESP_model.ply").c_str(), pts, npts) != 0) {
                rc = fail(err, err_len, HPM_EIO, "cannot write %s/ACMM_NESP_model.ply", out.c_str());
            }
            if (num_points) {
                *num_points = npts;
            }
        }
        hpm_fusion_free(pts);
    }
    hpm_pool_destroy(pool);
    return rc;
}
