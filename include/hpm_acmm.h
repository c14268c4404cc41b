/*
 * hpm_acmm.h — the ACMM+NESP coarse-to-fine schedule
 * (ACMM+NESP/windows/main.cpp:35-378) on the device, with the ACMM variant
 * of the PatchMatch kernels (hpm_params.acmm, hpm_types.h).
 *
 * Per problem (one reference view), ComputeMultiScaleSettings
 * (main.cpp:35-72): max_size = min(max(W, H), max_image_size), k = the
 * number of halvings that bring it to scale_bound (1000) or below. Levels run
 * from the coarsest (max over problems of k) to the full size; at a level a
 * view's image is downscaled to cur = max_size / 2^k_left (factor min(cur /
 * W, cur / H), ACMM.cpp:497-523), and
 *   - the first level runs a plain RunPatchMatch per view, later levels a
 *     hierarchy RunPatchMatch from the view's final maps at the previous
 *     level (JointBilateralUpsampling + SetHierarchyParams, main.cpp:118-138,
 *     338-351);
 *   - then geom_iterations (2) geometric-consistency passes, the first
 *     reading the level's first-pass depth maps, later ones the latest
 *     geometric maps in view order (main.cpp:325-372).
 * Views of one pass that do not depend on each other run concurrently on
 * host workers, as in hpm_pipeline_run.
 */
#ifndef HPM_ACMM_H_
#define HPM_ACMM_H_

#include <stdint.h>

#include "hpm_capi.h"
#include "hpm_pipeline.h"
#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpm_acmm_options {
    hpm_params params;        /* acmm = 1; max_image_size, seeds */
    int32_t scale_bound;      /* size_bound of ComputeMultiScaleSettings (main.cpp:38), 1000 */
    int32_t geom_iterations;  /* geometric passes per level (main.cpp:316), 2 */
    int32_t workers;          /* host workers, each with its own handle and stream (12) */
    int32_t verbose;
} hpm_acmm_options;

void hpm_acmm_default(hpm_acmm_options* opt);

/* main() of ACMM+NESP up to the fusion (main.cpp:302-374) on in-memory
 * views: images[v] is cams[v].width x cams[v].height row-major grey (full
 * size). Problem i has reference view ref_ids[i] and sources
 * src_ids[i * max_src + j], j < num_src[i]; every view used as a source
 * must be the reference of a problem (its level sizes come from that
 * problem, ACMM.cpp:497-500), and every view must change size from one
 * level to the next. *out_pool receives the pool of the last level, whose
 * state holds every view's final (world normal, depth) and costs and whose
 * depth maps hold the level's first-pass (0) and geometric (1) depths;
 * destroy it with hpm_pool_destroy. stats (may be NULL) sums the levels.
 * Returns 0, or a negative code with a message in err. */
int hpm_acmm_run(int device, int num_views, const hpm_camera* cams, const float* const* images, int num_problems,
                 const int* ref_ids, const int* num_src, const int* src_ids, int max_src,
                 const hpm_acmm_options* opt, hpm_pool** out_pool, hpm_pipeline_stats* stats, char* err,
                 int err_len);

/* main() (main.cpp:302-378) on a folder: pair.txt (sources with score > 0),
 * images/%08d.jpg or .pgm, cams/%08d_cam.txt; writes
 * ACMM_NESP/2333_%08d/{depths.dmb, depths_geom.dmb, normals.dmb, costs.dmb}
 * of the last level and, with fuse set, RunFusion's
 * ACMM_NESP/ACMM_NESP_model.ply (main.cpp:164-300). *num_points (may be NULL)
 * receives the fused point count. */
int hpm_acmm_run_folder(const char* dense_folder, const hpm_acmm_options* opt, int fuse, int64_t* num_points,
                        hpm_pipeline_stats* stats, char* err, int err_len);

#ifdef __cplusplus
}
#endif

#endif /* HPM_ACMM_H_ */
