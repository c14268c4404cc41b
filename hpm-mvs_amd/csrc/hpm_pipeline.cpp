/*
 * hpm_pipeline.cpp — ProcessProblem and the pass schedule of main()
 * (HPM-MVS/main.cpp:35-499, 636-674) on a device-resident pool
 * (include/hpm_pipeline.h).
 */
#include "hpm_pipeline.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hpm_batch.h"
#include "hpm_internal.h"
#include "hpm_io.h"
#include "hpm_prior.h"
#include "hpm_unet.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0)
{
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

struct Ctx {
    hpm_pool* pool;
    const hpm_pipeline_options* opt;
    hpm_pipeline_stats* st;
    std::string err;
    int threads;      /* host threads for the KNN supplement of one worker */
    std::mutex mu;    /* guards st and err across phase-A workers */
    hpm_lprior* lprior = nullptr; /* learned prior, when a weight file is given */
};

int fail(Ctx& c, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> lk(c.mu);
    c.err = buf;
    return HPM_EINVAL;
}

int check(Ctx& c, hpm_handle* h, int rc, const char* what)
{
    if (rc < 0) {
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.err.empty()) {
            c.err = std::string(what) + ": " + h->err;
        }
    }
    return rc;
}

void add_ms(Ctx& c, double* field, double ms)
{
    std::lock_guard<std::mutex> lk(c.mu);
    *field += ms;
}

/* one hpm_run_batch call over n problems */
void account_batch(Ctx& c, const hpm_batch_stats& bs, int n)
{
    std::lock_guard<std::mutex> lk(c.mu);
    c.st->ms_device += bs.ms_device;
    c.st->n_eval += bs.n_eval;
    c.st->runs += n;
    c.st->batches += 1;
}

void account(Ctx& c, const hpm_handle* h)
{
    std::lock_guard<std::mutex> lk(c.mu);
    c.st->ms_device += h->stats.ms_total;
    c.st->ms_propagate += h->stats.ms_propagate;
    c.st->n_eval += h->stats.n_eval;
    c.st->n_eval_propagate += h->stats.n_eval_propagate;
    c.st->runs += 1;
    c.st->propagate_launches += 6;
}

/* host maps of the current result: depth, world normal (3 floats), cost */
struct Maps {
    int W = 0, H = 0;
    std::vector<float> depth, normal, cost;
};

void maps_from_handle(const hpm_handle* h, Maps* m)
{
    m->W = h->cams[0].width;
    m->H = h->cams[0].height;
    const size_t np = (size_t)m->W * m->H;
    m->depth.resize(np);
    m->normal.resize(np * 3);
    m->cost.assign(h->h_cost.begin(), h->h_cost.end());
    for (size_t i = 0; i < np; ++i) {
        m->normal[3 * i + 0] = h->h_hyp[4 * i + 0];
        m->normal[3 * i + 1] = h->h_hyp[4 * i + 1];
        m->normal[3 * i + 2] = h->h_hyp[4 * i + 2];
        m->depth[i] = h->h_hyp[4 * i + 3];
    }
}

/* TransformNormal (HPM.cpp:872-880): camera -> world */
void to_world(const hpm_camera& c, const float n[4], float out[3])
{
    out[0] = c.R[0] * n[0] + c.R[3] * n[1] + c.R[6] * n[2];
    out[1] = c.R[1] * n[0] + c.R[4] * n[1] + c.R[7] * n[2];
    out[2] = c.R[2] * n[0] + c.R[5] * n[1] + c.R[8] * n[2];
}

/* Support points, Delaunay triangles, rasterised triangle mask, triangle
 * planes and the KNN supplement at one scale (main.cpp:84-121,
 * 238-273, 387-424). mask: idx+1 inside triangles, -233 for KNN planes. */
struct PriorScale {
    int w = 0, h = 0;
    std::vector<float> mask;
    std::vector<float> tri_planes; /* float4 per triangle */
    std::vector<float> knn_planes; /* float4 per pixel */
};

void build_prior_scale(Ctx& c, const hpm_camera& ref, const float* depth, const float* cost, int w, int h,
                       float factor, float dmin, float dmax, PriorScale* ps)
{
    ps->w = w;
    ps->h = h;
    auto t0 = clk::now();
    std::vector<int32_t> pts((size_t)(w / 5 + 1) * (h / 5 + 1) * 2);
    const int npts = hpm_support_points(cost, w, h, pts.data(), (int)(pts.size() / 2));
    pts.resize((size_t)npts * 2);
    std::vector<int32_t> tris;
    if (npts >= 3) {
        /* a triangulation of n points has fewer than 2n triangles */
        tris.resize((size_t)npts * 2 * 3);
        const int nt = hpm_delaunay(pts.data(), npts, tris.data(), npts * 2);
        tris.resize((size_t)std::max(nt, 0) * 3);
    }
    const int ntris = (int)(tris.size() / 3);
    auto t1 = clk::now();
    ps->mask.assign((size_t)w * h, 0.0f);
    hpm_rasterize_triangles(pts.data(), tris.data(), ntris, ps->mask.data(), w, h);
    auto t2 = clk::now();
    ps->tri_planes.assign((size_t)ntris * 4, 0.0f);
    for (int t = 0; t < ntris; ++t) {
        const int32_t p[6] = {pts[2 * tris[3 * t]], pts[2 * tris[3 * t] + 1], pts[2 * tris[3 * t + 1]],
                              pts[2 * tris[3 * t + 1] + 1], pts[2 * tris[3 * t + 2]], pts[2 * tris[3 * t + 2] + 1]};
        hpm_plane_from_points(&ref, p, depth, w, factor, &ps->tri_planes[4 * (size_t)t]);
    }
    auto t3 = clk::now();
    ps->knn_planes.assign((size_t)w * h * 4, 0.0f);
    hpm_knn_supplement(&ref, pts.data(), npts, depth, w, h, factor, c.opt->params.knn_k, dmin, dmax, ps->mask.data(),
                       ps->knn_planes.data(), c.threads);
    if (c.opt->verbose >= 2) {
        auto ms = [](clk::time_point a, clk::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        fprintf(stderr,
                "prior %dx%d: %d points %d triangles | support+delaunay %.1f ms, raster %.1f ms, planes %.1f ms, "
                "knn %.1f ms (%d threads)\n",
                w, h, npts, ntris, ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, clk::now()), c.threads);
    }
}

/* ProcessProblem (main.cpp:35-499) for phase A. */
/* ProcessProblem up to the plain pass: the handle set up for it */
int phase_a_setup(Ctx& c, hpm_handle* h, int ref, const int* srcs, int nsrc)
{
    hpm_params p = c.opt->params;
    p.planar_prior = 0;
    p.geom_consistency = 0;
    p.multi_geometry = 0;
    h->params = p;
    int rc = check(c, h, hpm_input_init_pool(h, c.pool, ref, srcs, nsrc), "input init");
    if (rc < 0) return rc;
    rc = check(c, h, hpm_device_init(h, nullptr), "device init");
    if (rc < 0) return rc;
    h->call_index = (uint32_t)c.opt->stream_base;
    if (const hpm_pool* cp = c.opt->hierarchy_from) {
        /* ACMM: the view's final maps at the coarser scale (normals.dmb,
         * depths_geom.dmb there; ACMM.cpp:660-717) */
        if (ref >= cp->n || !cp->have_state[(size_t)ref]) {
            return fail(c, "view %d: no state in the coarser pool", ref);
        }
        rc = check(c, h,
                   hpm_set_hierarchy_internal(h, cp->d_state_hyp + 4 * cp->off[(size_t)ref], cp->cams[(size_t)ref].width,
                                              cp->cams[(size_t)ref].height, true),
                   "hierarchy input");
        if (rc < 0) return rc;
    }
    return HPM_OK;
}

/* the learned prior on a finished plain pass (network + judge); *use tells
 * whether the view keeps it for one prior pass, else the handle is left for
 * the planar prior */
int learned_apply(Ctx& c, hpm_handle* h, int ref, bool* use)
{
    auto t0 = clk::now();
    float kept = 0.0f;
    int rc = check(c, h, hpm_lprior_apply(c.lprior, h, &kept), "learned prior");
    if (rc < 0) return rc;
    add_ms(c, &c.st->ms_prior_device, ms_since(t0));
    *use = kept >= c.opt->learned_min_accept;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        c.st->learned_accept += kept;
        if (*use) {
            c.st->learned_views += 1;
        } else {
            c.st->learned_fallbacks += 1;
        }
    }
    if (c.opt->verbose >= 2) {
        fprintf(stderr, "view %d learned prior: %.3f of the pixels kept\n", ref, kept);
    }
    if (!*use) {
        /* fallback: the planar prior, from the plain pass */
        h->params.planar_prior = 0;
        h->have_prior = false;
    }
    return HPM_OK;
}

/* ProcessProblem after the plain pass: the planar prior passes (device or
 * host construction), then the maps go to the pool */
int phase_a_planar(Ctx& c, hpm_handle* h, int ref)
{
    int rc = HPM_OK;
    const hpm_camera ref_cam = h->cams[0];
    const int W = ref_cam.width, H = ref_cam.height;
    const float dmin = h->params.depth_min, dmax = h->params.depth_max;
    if (c.opt->planar_prior && c.opt->prior_on_device) {
        hpm_set_planar_prior(h); /* SetPlanarPriorParams (HPM.cpp:477-480) */
        int k = 0;
        for (int max_size = std::max(W, H); max_size > c.opt->scale_bound; max_size /= 2) {
            ++k;
        }
        for (int i = 2 - k; i <= 2; ++i) {
            const float f = i == 0 ? 0.25f : (i == 1 ? 0.5f : 1.0f);
            auto t0 = clk::now();
            double host_ms = 0.0;
            rc = check(c, h, hpm_build_planar_prior(h, f, c.opt->params.knn_k, &host_ms), "planar prior");
            if (rc < 0) return rc;
            add_ms(c, &c.st->ms_prior_host, host_ms);
            add_ms(c, &c.st->ms_prior_device, ms_since(t0) - host_ms);
            if (c.opt->verbose >= 2) {
                fprintf(stderr, "view %d prior %.2fx on the device: %.1f ms (host %.1f ms)\n", ref, f, ms_since(t0),
                        host_ms);
            }
            rc = check(c, h, hpm_run_internal(h, false), "RunPatchMatch (prior)");
            if (rc < 0) return rc;
            account(c, h);
        }
        return check(c, h, hpm_store_to_pool(h, 0), "store");
    }
    if (c.opt->planar_prior) {
        hpm_set_planar_prior(h); /* SetPlanarPriorParams (HPM.cpp:477-480) */
        Maps m;
        maps_from_handle(h, &m);
        /* ComputingMultiScaleSettings (HPM.cpp:851-860) with its bound */
        int k = 0;
        for (int max_size = std::max(W, H); max_size > c.opt->scale_bound; max_size /= 2) {
            ++k;
        }
        for (int i = 2 - k; i <= 2; ++i) {
            auto t0 = clk::now();
            if (i == 0 || i == 1) {
                const float f = i == 0 ? 0.25f : 0.5f;
                const int ws = (int)lround(W * (double)f), hs = (int)lround(H * (double)f);
                std::vector<float> ds((size_t)ws * hs), cs((size_t)ws * hs), ns((size_t)ws * hs * 3);
                hpm_resize_bilinear_c(m.depth.data(), W, H, 1, ds.data(), ws, hs);
                hpm_resize_bilinear_c(m.cost.data(), W, H, 1, cs.data(), ws, hs);
                hpm_resize_bilinear_c(m.normal.data(), W, H, 3, ns.data(), ws, hs);
                PriorScale ps;
                build_prior_scale(c, ref_cam, ds.data(), cs.data(), ws, hs, f, dmin, dmax, &ps);
                /* low-scale prior maps (main.cpp:124-162, 277-314), world normals */
                std::vector<float> src_nd((size_t)ws * hs * 4, 0.0f);
                for (int x = 0; x < ws; ++x) {
                    for (int y = 0; y < hs; ++y) {
                        const size_t q = (size_t)y * ws + x;
                        float& mk = ps.mask[q];
                        const float* plane = nullptr;
                        bool ok = false;
                        if (mk > 0) {
                            plane = &ps.tri_planes[4 * (size_t)((int)mk - 1)];
                            const float d = hpm_depth_from_plane_factor(&ref_cam, plane, x, y, f);
                            ok = d <= dmax * 1.2f && d >= dmin * 0.6f;
                            if (ok) src_nd[4 * q + 3] = d;
                        } else if (mk < -200) {
                            plane = &ps.knn_planes[4 * q];
                            const float d = hpm_depth_from_plane_factor(&ref_cam, plane, x, y, f);
                            ok = d <= dmax && d >= dmin;
                            if (ok) src_nd[4 * q + 3] = d;
                        }
                        if (ok) {
                            to_world(ref_cam, plane, &src_nd[4 * q]);
                        } else {
                            mk = 0;
                        }
                    }
                }
                add_ms(c, &c.st->ms_prior_host, ms_since(t0));
                if (c.opt->verbose >= 2) {
                    fprintf(stderr, "view %d prior %.2fx: host %.1f ms\n", ref, f, ms_since(t0));
                }
                auto t1 = clk::now();
                rc = check(c, h, hpm_jbu_prior(h, src_nd.data(), ws, hs, f), "JBU prior");
                if (rc < 0) return rc;
                if (rc == 1) {
                    /* integer scale 1: JointBilateralUpsampling returns without
                     * output (HPM.cpp:1219-1222), the upsampled maps stay 0 and
                     * every mask is 0 */
                    const size_t np = (size_t)W * H;
                    std::vector<float> zp(np * 4, 0.0f);
                    std::vector<uint32_t> zm(np, 0u);
                    rc = check(c, h, hpm_upload_prior(h, zp.data(), zm.data()), "prior upload");
                    if (rc < 0) return rc;
                }
                add_ms(c, &c.st->ms_prior_device, ms_since(t1));
            } else {
                PriorScale ps;
                build_prior_scale(c, ref_cam, m.depth.data(), m.cost.data(), W, H, 1.0f, dmin, dmax, &ps);
                /* range checks of main.cpp:427-454 and the upload of
                 * CudaPlanarPriorInitializationSupplement (HPM.cpp:1099-1123):
                 * KNN pixels count as active (Q12) */
                const size_t np = (size_t)W * H;
                std::vector<float> planes(np * 4, 0.0f);
                std::vector<uint32_t> mask(np, 0u);
                for (int x = 0; x < W; ++x) {
                    for (int y = 0; y < H; ++y) {
                        const size_t q = (size_t)y * W + x;
                        const float mk = ps.mask[q];
                        const float* plane = nullptr;
                        if (mk > 0) {
                            plane = &ps.tri_planes[4 * (size_t)((int)mk - 1)];
                        } else if (mk < -200) {
                            plane = &ps.knn_planes[4 * q];
                        }
                        if (plane) {
                            const float d = hpm_depth_from_plane_factor(&ref_cam, plane, x, y, 1.0f);
                            if (d <= dmax && d >= dmin) {
                                memcpy(&planes[4 * q], plane, 16);
                                mask[q] = 1u;
                            }
                        }
                    }
                }
                add_ms(c, &c.st->ms_prior_host, ms_since(t0));
                if (c.opt->verbose >= 2) {
                    fprintf(stderr, "view %d prior 1x: host %.1f ms\n", ref, ms_since(t0));
                }
                auto t1 = clk::now();
                rc = check(c, h, hpm_upload_prior(h, planes.data(), mask.data()), "prior upload");
                if (rc < 0) return rc;
                add_ms(c, &c.st->ms_prior_device, ms_since(t1));
            }
            rc = check(c, h, hpm_run_internal(h, true), "RunPatchMatch (prior)");
            if (rc < 0) return rc;
            account(c, h);
            maps_from_handle(h, &m);
        }
    }
    return check(c, h, hpm_store_to_pool(h, 0), "store");
}

int process_phase_a(Ctx& c, hpm_handle* h, int ref, const int* srcs, int nsrc)
{
    int rc = phase_a_setup(c, h, ref, srcs, nsrc);
    if (rc < 0) return rc;
    /* the host builds the planar prior from the maps it reads back */
    const bool host_prior = c.opt->planar_prior != 0 && !c.opt->prior_on_device;
    rc = check(c, h, hpm_run_internal(h, host_prior), "RunPatchMatch");
    if (rc < 0) return rc;
    account(c, h);
    if (c.opt->planar_prior && c.lprior) {
        bool use = false;
        rc = learned_apply(c, h, ref, &use);
        if (rc < 0) return rc;
        if (use) {
            rc = check(c, h, hpm_run_internal(h, false), "RunPatchMatch (learned prior)");
            if (rc < 0) return rc;
            account(c, h);
            return check(c, h, hpm_store_to_pool(h, 0), "store");
        }
    }
    return phase_a_planar(c, h, ref);
}

/* ProcessProblem with geometric consistency (main.cpp:44-51, 481-490) up
 * to its RunPatchMatch */
int phase_b_setup(Ctx& c, hpm_handle* h, int ref, const int* srcs, int nsrc, int pass)
{
    const bool multi = pass > 0;
    hpm_params p = c.opt->params;
    p.planar_prior = 0;
    p.geom_consistency = 0;
    p.multi_geometry = 0;
    h->params = p;
    hpm_set_geom_consistency(h, multi ? 1 : 0);
    int rc = check(c, h, hpm_input_init_pool(h, c.pool, ref, srcs, nsrc), "input init");
    if (rc < 0) return rc;
    rc = check(c, h, hpm_device_init(h, nullptr), "device init");
    if (rc < 0) return rc;
    rc = check(c, h, hpm_load_from_pool(h), "reload");
    if (rc < 0) return rc;
    h->call_index = (uint32_t)c.opt->stream_base + 8u + (uint32_t)pass;
    return HPM_OK;
}

int process_phase_b(Ctx& c, hpm_handle* h, int ref, const int* srcs, int nsrc, int pass)
{
    int rc = phase_b_setup(c, h, ref, srcs, nsrc, pass);
    if (rc < 0) return rc;
    rc = check(c, h, hpm_run_internal(h, false), "RunPatchMatch (geometric)");
    if (rc < 0) return rc;
    account(c, h);
    return check(c, h, hpm_store_to_pool(h, 1), "store");
}

/* a problem handle for a schedule over pool: a spare one left by an earlier
 * schedule (its stream and buffers kept), else a new one */
hpm_handle* take_handle(hpm_pool* pool)
{
    {
        std::lock_guard<std::mutex> lk(pool->spare_mu);
        if (!pool->spare.empty()) {
            hpm_handle* h = pool->spare.back();
            pool->spare.pop_back();
            return h;
        }
    }
    return hpm_create(pool->device);
}

/* gives a finished schedule's handle back to the pool; beyond 64 spare
 * handles it is destroyed */
void give_handle(hpm_pool* pool, hpm_handle* h)
{
    if (!h) {
        return;
    }
    (void)hipStreamSynchronize(h->stream);
    h->err.clear();
    {
        std::lock_guard<std::mutex> lk(pool->spare_mu);
        if (pool->spare.size() < 64) {
            pool->spare.push_back(h);
            return;
        }
    }
    hpm_destroy(h);
}

/* the learned prior of a weight file: the pool's loaded copy while the file
 * (path, size, modification time) is unchanged, else loaded anew. The pool
 * keeps it; NULL if it cannot be loaded. */
hpm_lprior* pool_lprior(hpm_pool* pool, const char* path)
{
    struct stat sb;
    if (stat(path, &sb) != 0) {
        return nullptr;
    }
    char key[64];
    snprintf(key, sizeof(key), "|%lld|%lld.%09ld", (long long)sb.st_size, (long long)sb.st_mtim.tv_sec,
             (long)sb.st_mtim.tv_nsec);
    const std::string k = std::string(path) + key;
    if (pool->lprior && pool->lprior_key == k) {
        return pool->lprior;
    }
    hpm_lprior* lp = hpm_lprior_load(path, pool->device);
    if (!lp) {
        return nullptr;
    }
    hpm_lprior_destroy(pool->lprior);
    pool->lprior = lp;
    pool->lprior_key = k;
    return lp;
}

/* RunPatchMatch of the listed problems in one hpm_run_batch call */
int run_batch(Ctx& c, const std::vector<hpm_handle*>& hs, bool readback)
{
    if (hs.empty()) {
        return HPM_OK;
    }
    hpm_batch_stats bs;
    const int rc = hpm_run_batch(c.opt->batch, hs.data(), (int)hs.size(), readback ? 1 : 0, &bs);
    if (rc < 0) {
        return fail(c, "batch: %s", hpm_batch_last_error(c.opt->batch));
    }
    account_batch(c, bs, (int)hs.size());
    return HPM_OK;
}

} // namespace

extern "C" void hpm_pipeline_default(hpm_pipeline_options* opt)
{
    memset(opt, 0, sizeof(*opt));
    hpm_params_default(&opt->params);
    opt->planar_prior = 1;
    opt->geom_consistency = 1;
    opt->num_threads = 0;
    opt->scale_bound = 1000;
    opt->workers = 12;
    opt->learned_prior = nullptr;
    opt->learned_min_accept = 0.3f;
    opt->hierarchy_from = nullptr;
    opt->stream_base = 0;
    opt->prior_on_device = 1;
    opt->batch = nullptr;
}

extern "C" int hpm_pool_get_state(hpm_pool* pool, int view, float* hyp_world4, float* cost)
{
    if (!pool || view < 0 || view >= pool->n) {
        return HPM_EINVAL;
    }
    if (!pool->have_state[(size_t)view]) {
        return HPM_EINVAL;
    }
    const int64_t off = pool->off[(size_t)view];
    const size_t np = pool->images[(size_t)view].size();
    if ((hyp_world4 && hipMemcpyAsync(hyp_world4, pool->d_state_hyp + 4 * off, np * 16, hipMemcpyDeviceToHost,
                                      pool->stream) != hipSuccess) ||
        (cost && hipMemcpyAsync(cost, pool->d_state_cost + off, np * 4, hipMemcpyDeviceToHost, pool->stream) !=
                     hipSuccess) ||
        hipStreamSynchronize(pool->stream) != hipSuccess) {
        return HPM_EDEVICE;
    }
    return HPM_OK;
}

extern "C" int hpm_pool_set_state(hpm_pool* pool, int view, const float* hyp_world4, const float* cost)
{
    if (!pool || !hyp_world4 || view < 0 || view >= pool->n) {
        return HPM_EINVAL;
    }
    const int64_t off = pool->off[(size_t)view];
    const size_t np = pool->images[(size_t)view].size();
    std::vector<float> zeros;
    if (!cost) {
        zeros.assign(np, 0.0f);
        cost = zeros.data();
    }
    if (hipMemcpyAsync(pool->d_state_hyp + 4 * off, hyp_world4, np * 16, hipMemcpyHostToDevice, pool->stream) !=
            hipSuccess ||
        hipMemcpyAsync(pool->d_state_cost + off, cost, np * 4, hipMemcpyHostToDevice, pool->stream) != hipSuccess ||
        hipStreamSynchronize(pool->stream) != hipSuccess) {
        return HPM_EDEVICE;
    }
    pool->have_state[(size_t)view] = 1;
    return HPM_OK;
}

extern "C" int hpm_pipeline_run(hpm_pool* pool, int num_problems, const int* ref_ids, const int* num_src,
                                const int* src_ids, int max_src, const hpm_pipeline_options* opt,
                                hpm_pipeline_stats* stats, char* err, int err_len)
{
    hpm_pipeline_stats local;
    hpm_pipeline_stats* st = stats ? stats : &local;
    memset(st, 0, sizeof(*st));
    hpm_pipeline_options defaults;
    hpm_pipeline_default(&defaults);
    Ctx c;
    c.pool = pool;
    c.opt = opt ? opt : &defaults;
    c.st = st;
    c.threads = c.opt->num_threads > 0 ? c.opt->num_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    int rc = HPM_OK;
    auto t_all = clk::now();
    hpm_handle* h = nullptr;
    if (!pool || num_problems < 1 || !ref_ids || !num_src || !src_ids || max_src < 1) {
        rc = fail(c, "bad pipeline arguments");
    } else if (!(h = take_handle(pool))) {
        rc = fail(c, "cannot create a problem handle");
    } else if (c.opt->learned_prior && c.opt->planar_prior &&
               !(c.lprior = pool_lprior(pool, c.opt->learned_prior))) {
        rc = fail(c, "cannot load the learned prior %s", c.opt->learned_prior);
    }
    /* host workers, each with its own handle and stream: views that do not
     * depend on each other run concurrently, overlapping one view's host
     * prior construction with the device work of the others */
    const int workers = std::max(1, std::min(c.opt->workers, num_problems));
    std::vector<hpm_handle*> handles((size_t)workers, nullptr);
    handles[0] = h;
    /* the batched form runs its host threads over the problems' own handles */
    for (int w = 1; rc == HPM_OK && !c.opt->batch && w < workers; ++w) {
        handles[(size_t)w] = take_handle(pool);
        if (!handles[(size_t)w]) {
            rc = fail(c, "cannot create a problem handle");
        }
    }
    const int all_threads = c.threads;
    /* runs fn(handle, problem) for every problem on the workers */
    auto parallel = [&](auto&& fn) {
        std::atomic<int> next(0);
        std::atomic<int> worst(HPM_OK);
        auto work = [&](hpm_handle* wh) {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= num_problems || worst.load() < 0) {
                    return;
                }
                const int r = fn(wh, i);
                if (r < 0) {
                    worst.store(r);
                }
            }
        };
        c.threads = std::max(1, all_threads / workers);
        std::vector<std::thread> pool_threads;
        for (int w = 1; w < workers; ++w) {
            pool_threads.emplace_back(work, handles[(size_t)w]);
        }
        work(h);
        for (auto& t : pool_threads) {
            t.join();
        }
        c.threads = all_threads;
        return worst.load();
    };
    if (rc == HPM_OK && c.opt->batch && c.opt->hierarchy_from) {
        rc = fail(c, "the batched schedule runs HPM-MVS problems (no hierarchy)");
    }
    /* the batched form: one handle per problem, each pass whose views are
     * independent runs as one hpm_run_batch call over all of them */
    std::vector<hpm_handle*> ph;
    if (rc == HPM_OK && c.opt->batch) {
        for (int i = 0; i < num_problems && rc == HPM_OK; ++i) {
            ph.push_back(take_handle(pool));
            if (!ph.back()) {
                rc = fail(c, "cannot create a problem handle");
            }
        }
    }
    /* runs fn(handle of problem i, i) for every problem on the workers */
    auto each = [&](auto&& fn) {
        std::atomic<int> next(0);
        std::atomic<int> worst(HPM_OK);
        auto work = [&]() {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= num_problems || worst.load() < 0) {
                    return;
                }
                const int r = fn(ph[(size_t)i], i);
                if (r < 0) {
                    worst.store(r);
                }
            }
        };
        c.threads = std::max(1, all_threads / workers);
        std::vector<std::thread> ts;
        for (int w = 1; w < workers; ++w) {
            ts.emplace_back(work);
        }
        work();
        for (auto& t : ts) {
            t.join();
        }
        c.threads = all_threads;
        return worst.load();
    };
    auto srcs_of = [&](int i) { return src_ids + (size_t)i * max_src; };
    if (rc == HPM_OK && c.opt->batch) {
        auto t0 = clk::now();
        const bool host_prior = c.opt->planar_prior != 0 && !c.opt->prior_on_device;
        /* the problems' set-up (allocations, view tables) on the host threads */
        rc = each([&](hpm_handle* hh, int i) { return phase_a_setup(c, hh, ref_ids[i], srcs_of(i), num_src[i]); });
        if (rc >= 0) rc = run_batch(c, ph, host_prior);
        std::vector<char> use((size_t)num_problems, 0);
        if (rc >= 0 && c.opt->planar_prior && c.lprior) {
            rc = each([&](hpm_handle* hh, int i) {
                bool u = false;
                const int r = learned_apply(c, hh, ref_ids[i], &u);
                use[(size_t)i] = u ? 1 : 0;
                return r;
            });
            std::vector<hpm_handle*> learned;
            for (int i = 0; i < num_problems; ++i) {
                if (use[(size_t)i]) learned.push_back(ph[(size_t)i]);
            }
            if (rc >= 0) rc = run_batch(c, learned, false);
        }
        if (rc >= 0) {
            rc = each([&](hpm_handle* hh, int i) {
                if (use[(size_t)i]) {
                    return check(c, hh, hpm_store_to_pool(hh, 0), "store");
                }
                return phase_a_planar(c, hh, ref_ids[i]);
            });
        }
        st->ms_phase_a = ms_since(t0);
        if (rc >= 0 && c.opt->geom_consistency) {
            t0 = clk::now();
            for (int g = 0; g < c.opt->params.geom_iterations && rc >= 0; ++g) {
                if (g == 0) {
                    /* the first pass reads only phase-A maps: one batch */
                    rc = each([&](hpm_handle* hh, int i) {
                        return phase_b_setup(c, hh, ref_ids[i], srcs_of(i), num_src[i], g);
                    });
                    if (rc >= 0) rc = run_batch(c, ph, false);
                    for (int i = 0; i < num_problems && rc >= 0; ++i) {
                        rc = check(c, ph[(size_t)i], hpm_store_to_pool(ph[(size_t)i], 1), "store");
                    }
                } else {
                    /* later passes read maps of this pass: problem order */
                    for (int i = 0; i < num_problems && rc >= 0; ++i) {
                        rc = process_phase_b(c, ph[(size_t)i], ref_ids[i], srcs_of(i), num_src[i], g);
                    }
                }
            }
            st->ms_phase_b = ms_since(t0);
        }
    } else if (rc == HPM_OK) {
        /* phase A: views are independent (main.cpp:658-661) */
        auto t0 = clk::now();
        rc = parallel([&](hpm_handle* wh, int i) {
            if (c.opt->verbose) {
                fprintf(stderr, "phase A: view %d\n", ref_ids[i]);
            }
            return process_phase_a(c, wh, ref_ids[i], src_ids + (size_t)i * max_src, num_src[i]);
        });
        st->ms_phase_a = ms_since(t0);
    }
    if (rc >= 0 && c.opt->geom_consistency && !c.opt->batch) {
        auto t0 = clk::now();
        for (int g = 0; g < c.opt->params.geom_iterations && rc >= 0; ++g) {
            auto pass = [&](hpm_handle* wh, int i) {
                if (c.opt->verbose) {
                    fprintf(stderr, "geometric pass %d: view %d\n", g, ref_ids[i]);
                }
                return process_phase_b(c, wh, ref_ids[i], src_ids + (size_t)i * max_src, num_src[i], g);
            };
            if (g == 0) {
                /* the first pass reads only the phase-A depth maps and
                 * writes each view's own state: views are independent */
                rc = parallel(pass);
            } else {
                /* later passes read the geometric depth maps of views
                 * already updated in the same pass (main.cpp:662-674):
                 * problem order */
                for (int i = 0; i < num_problems && rc >= 0; ++i) {
                    rc = pass(h, i);
                }
            }
        }
        st->ms_phase_b = ms_since(t0);
    }
    for (int w = 1; w < workers; ++w) {
        give_handle(pool, handles[(size_t)w]);
    }
    for (hpm_handle* x : ph) {
        give_handle(pool, x);
    }
    give_handle(pool, h);
    if (st->learned_views + st->learned_fallbacks > 0) {
        st->learned_accept /= (double)(st->learned_views + st->learned_fallbacks);
    }
    st->ms_wall = ms_since(t_all);
    if (rc < 0 && err && err_len > 0) {
        snprintf(err, (size_t)err_len, "%s", c.err.c_str());
    }
    return rc < 0 ? rc : HPM_OK;
}

extern "C" int hpm_pipeline_run_folder(const char* dense_folder, const hpm_pipeline_options* opt,
                                       hpm_pipeline_stats* stats, char* err, int err_len)
{
    auto say = [&](const char* msg) {
        if (err && err_len > 0) {
            snprintf(err, (size_t)err_len, "%s", msg);
        }
    };
    if (!dense_folder) {
        say("null folder");
        return HPM_EINVAL;
    }
    const std::string base(dense_folder);
    int np = 0;
    if (hpm_read_pair((base + "/pair.txt").c_str(), &np, nullptr, nullptr, nullptr, 0) != 0 || np < 1) {
        say("cannot read pair.txt");
        return HPM_EIO;
    }
    std::vector<int> refs((size_t)np), ns((size_t)np), srcs((size_t)np * HPM_MAX_SOURCES);
    hpm_read_pair((base + "/pair.txt").c_str(), &np, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES);
    int maxid = 0;
    for (int i = 0; i < np; ++i) {
        maxid = std::max(maxid, refs[(size_t)i]);
        for (int j = 0; j < ns[(size_t)i]; ++j) {
            maxid = std::max(maxid, srcs[(size_t)i * HPM_MAX_SOURCES + j]);
        }
    }
    /* the pool holds views 0..maxid by image id */
    std::vector<hpm_camera> cams((size_t)maxid + 1);
    std::vector<std::vector<float>> imgs((size_t)maxid + 1);
    for (int id = 0; id <= maxid; ++id) {
        char path[4096];
        int w = 0, hh = 0;
        snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder, id);
        if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, id);
            if (hpm_read_image_grey(path, nullptr, &w, &hh) != 0) {
                say("cannot read an image (JPEG or 8-bit PGM expected)");
                return HPM_EIO;
            }
        }
        imgs[(size_t)id].resize((size_t)w * hh);
        hpm_read_image_grey(path, imgs[(size_t)id].data(), &w, &hh);
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, id);
        if (hpm_read_camera(path, &cams[(size_t)id]) != 0) {
            say("cannot read a camera file");
            return HPM_EIO;
        }
        cams[(size_t)id].width = w;
        cams[(size_t)id].height = hh;
    }
    std::vector<const float*> ip;
    for (auto& v : imgs) ip.push_back(v.data());
    hpm_pipeline_options defaults;
    hpm_pipeline_default(&defaults);
    const hpm_pipeline_options* o = opt ? opt : &defaults;
    hpm_pool* pool = hpm_pool_create(0, maxid + 1, cams.data(), ip.data(), o->params.max_image_size);
    if (!pool) {
        say("cannot create the device pool");
        return HPM_EDEVICE;
    }
    int rc = hpm_pipeline_run(pool, np, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES, o, stats, err, err_len);
    if (rc == HPM_OK) {
        mkdir((base + "/HPM").c_str(), 0755);
        for (int i = 0; i < np && rc == HPM_OK; ++i) {
            const int v = refs[(size_t)i];
            const size_t n = pool->images[(size_t)v].size();
            const int W = pool->cams[(size_t)v].width, H = pool->cams[(size_t)v].height;
            std::vector<float> hyp(n * 4), cost(n), nrm(n * 3), d_plain(n), d_geom(n);
            hpm_pool_get_state(pool, v, hyp.data(), cost.data());
            hpm_pool_get_depth(pool, v, 0, d_plain.data());
            hpm_pool_get_depth(pool, v, 1, d_geom.data());
            for (size_t k = 0; k < n; ++k) {
                nrm[3 * k + 0] = hyp[4 * k + 0];
                nrm[3 * k + 1] = hyp[4 * k + 1];
                nrm[3 * k + 2] = hyp[4 * k + 2];
            }
            char dir[4096];
            snprintf(dir, sizeof(dir), "%s/HPM/2333_%08d", dense_folder, v);
            mkdir(dir, 0755);
            const std::string d(dir);
            if (hpm_write_dmb((d + "/depths.dmb").c_str(), d_plain.data(), W, H, 1) != 0 ||
                hpm_write_dmb((d + "/normals.dmb").c_str(), nrm.data(), W, H, 3) != 0 ||
                hpm_write_dmb((d + "/costs.dmb").c_str(), cost.data(), W, H, 1) != 0 ||
                (o->geom_consistency && hpm_write_dmb((d + "/depths_geom.dmb").c_str(), d_geom.data(), W, H, 1) != 0)) {
                say("cannot write results");
                rc = HPM_EIO;
            }
        }
    }
    hpm_pool_destroy(pool);
    return rc;
}
