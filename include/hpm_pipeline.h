/*
 * hpm_pipeline.h — the per-scene pass schedule of HPM-MVS (main.cpp:35-499,
 * 636-674) over a device-resident pool:
 *   phase A, every view: plain RunPatchMatch, then (planar prior) one prior
 *     RunPatchMatch per scale i = 2-k .. 2 of ComputingMultiScaleSettings:
 *     low scales build the prior at 0.25x / 0.5x and upsample it with JBU,
 *     the full scale builds it at 1x (triangles + KNN supplement);
 *   phase B: geom_iterations geometric-consistency passes over all views,
 *     the first reading the phase-A depth maps, later ones the latest
 *     geometric depth maps (views in order, as the reference runs them).
 * Phase A and the first geometric pass run their views concurrently on
 * `workers` host threads (views are independent there); later geometric
 * passes run in problem order, each view reading the geometric depth maps
 * its predecessors wrote in the same pass.
 *
 * Random streams: RunPatchMatch call c of phase A uses Philox stream
 * c (0..k); geometric pass g uses stream 8 + g, so no two passes of a view
 * share random numbers (the reference reseeds from clock64() per launch).
 */
#ifndef HPM_PIPELINE_H_
#define HPM_PIPELINE_H_

#include <stdint.h>

#include "hpm_batch.h"
#include "hpm_capi.h"
#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpm_pipeline_options {
    hpm_params params;          /* seeds, knn_k and geom_iterations are read from here */
    int32_t planar_prior;       /* phase A with the multi-scale planar prior (main.cpp:654) */
    int32_t geom_consistency;   /* run phase B (main.cpp:662-674) */
    int32_t num_threads;        /* host threads for the prior construction (<= 0: all) */
    int32_t verbose;
    int32_t scale_bound;        /* size_bound of ComputingMultiScaleSettings (HPM.cpp:852), 1000 */
    int32_t workers;            /* host workers of the independent passes, each with its own handle and stream (12) */
    /* learned prior (no reference counterpart): weight file of the U-Net
     * (hpm_unet.h); when set, phase A runs the plain pass, the network and
     * its judge, and one prior pass instead of the planar prior scales, and
     * falls back to the planar prior for a view where fewer than
     * learned_min_accept of the pixels keep the network's proposal */
    const char* learned_prior;  /* NULL: planar prior */
    float learned_min_accept;   /* 0.3 */
    /* ACMM hierarchy (params.acmm, no planar prior): when set, phase A runs
     * each view once from its final state in this coarser pool
     * (hpm_set_hierarchy_input) instead of a plain pass */
    hpm_pool* hierarchy_from;
    int32_t stream_base;        /* added to every RunPatchMatch call's Philox stream (0) */
    int32_t prior_on_device;    /* build the planar prior with hpm_build_planar_prior (1), or on the host (0) */
    /* the batched form (no reference counterpart): when set, every pass
     * whose views are independent (the plain pass, the learned-prior pass,
     * the first geometric pass) runs as one hpm_run_batch call over all
     * problems on this caller-owned object, which keeps the tuned and
     * captured schedules across runs; the planar prior passes stay per
     * view on the workers. NULL: every pass per view on the workers. The
     * maps equal those of the per-view schedule bit for bit. */
    hpm_batch* batch;
} hpm_pipeline_options;

typedef struct hpm_pipeline_stats {
    double ms_wall;             /* whole schedule, host clock */
    double ms_phase_a;
    double ms_phase_b;
    double ms_prior_host;       /* support points, Delaunay, rasterisation, plane fits, KNN */
    double ms_prior_device;     /* JBU and prior uploads */
    double ms_device;           /* device time of all RunPatchMatch calls (hipEvents) */
    uint64_t n_eval;            /* NCC evaluations of all RunPatchMatch calls */
    uint64_t n_eval_propagate;
    double ms_propagate;        /* device time of the propagation launches */
    int32_t runs;               /* RunPatchMatch calls */
    int32_t propagate_launches;
    int32_t learned_views;      /* views whose prior pass used the learned prior */
    int32_t learned_fallbacks;  /* views that fell back to the planar prior */
    double learned_accept;      /* mean kept fraction over the views that ran the network */
    int32_t batches;            /* hpm_run_batch calls of the batched form */
} hpm_pipeline_stats;

void hpm_pipeline_default(hpm_pipeline_options* opt);

/* Runs the schedule for num_problems problems on the pool: problem i has
 * reference view ref_ids[i] and sources src_ids[i * max_src + j],
 * j < num_src[i]. Final maps stay in the pool (hpm_pool_get_state). err
 * receives a message on failure. */
int hpm_pipeline_run(hpm_pool* pool, int num_problems, const int* ref_ids, const int* num_src, const int* src_ids,
                     int max_src, const hpm_pipeline_options* opt, hpm_pipeline_stats* stats, char* err,
                     int err_len);

/* Last (world normal xyz, depth) hypotheses and costs of a view. */
int hpm_pool_get_state(hpm_pool* pool, int view, float* hyp_world4, float* cost);
/* Sets them (for instance from normals.dmb and a depth map, to fuse maps
 * written by an earlier run); cost may be NULL (zeros). */
int hpm_pool_set_state(hpm_pool* pool, int view, const float* hyp_world4, const float* cost);

/* main() (main.cpp:636-677) on a folder: pair.txt, images/, cams/; writes
 * HPM/2333_%08d/{depths.dmb, depths_geom.dmb, normals.dmb, costs.dmb}. */
int hpm_pipeline_run_folder(const char* dense_folder, const hpm_pipeline_options* opt, hpm_pipeline_stats* stats,
                            char* err, int err_len);

#ifdef __cplusplus
}
#endif

#endif /* HPM_PIPELINE_H_ */
