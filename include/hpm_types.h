/*
 * hpm_types.h — plain-old-data types shared by the C-ABI, the host runtime,
 * the HIP kernels and the CPU oracle.
 *
 * hpm_camera keeps the field order and meaning of the reference's `Camera`
 * (HPM-MVS/main.h:44-52): K and R row-major, world->camera x_c = R*X + t.
 * hpm_params keeps the defaults of `PatchMatchParams` (HPM-MVS/HPM.h:29-46)
 * plus the knobs that the reference hard-codes and that the workloads here
 * change (KNN K, geometric-consistency pass count, the RNG seed).
 */
#ifndef HPM_TYPES_H_
#define HPM_TYPES_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Maximum number of images per problem (1 reference + up to 32 sources).
 * The reference keeps per-pixel view masks in a 32-bit word
 * (HPM-MVS/HPM.cu:55-63) and cost vectors in float[32] (HPM.cu:402). */
#define HPM_MAX_SOURCES 32
#define HPM_MAX_IMAGES (HPM_MAX_SOURCES + 1)

/* HPM-MVS/main.h:44-52 */
typedef struct hpm_camera {
    float K[9];
    float R[9];
    float t[3];
    int32_t height;
    int32_t width;
    float depth_min;
    float depth_max;
} hpm_camera;

/* HPM-MVS/HPM.h:29-46 (same defaults), plus explicit settings. */
typedef struct hpm_params {
    int32_t max_iterations;   /* stored, like the reference; RunPatchMatch always runs 3 (HPM.cu:1933) */
    int32_t patch_size;       /* 11 */
    int32_t num_images;       /* 1 + number of sources */
    int32_t max_image_size;   /* 3200 (HPM.h:33) */
    int32_t radius_increment; /* 2 */
    float sigma_spatial;      /* 5 */
    float sigma_color;        /* 3 */
    int32_t top_k;            /* 4 */
    float baseline;           /* 0.54 (unused on the path) */
    float depth_min;          /* set by input init: 0.6 * cam depth_min (HPM.cpp:553) */
    float depth_max;          /* set by input init: 1.2 * cam depth_max (HPM.cpp:554) */
    float disparity_min;
    float disparity_max;
    int32_t geom_consistency;
    int32_t multi_geometry;
    int32_t planar_prior;
    /* settings the reference hard-codes */
    int32_t knn_k;            /* 6 (HPM.cpp:962) */
    int32_t geom_iterations;  /* 2 (main.cpp:656) */
    uint32_t seed_lo;         /* Philox key; the reference seeds from clock64() (HPM.cu:521) */
    uint32_t seed_hi;
    /* The ACMM+NESP variant (ACMM+NESP/windows/ACMM.cu, PatchMatchParams at
     * ACMM.h:21-45); both 0 for HPM-MVS. acmm: the geometric cost is capped at
     * 3 (ACMM.cu:496) and weighted 0.2 in the re-score and the refinement
     * (ACMM.cu:1596, 651), adoption has no NaN fallback (ACMM.cu:1607-1616),
     * and RunPatchMatch runs max_iterations (1..3) red/black iterations
     * (ACMM.cu:1836), 2 in geometric passes (ACMM.cpp:434-441). hierarchy
     * (SetHierarchyParams, ACMM.cpp:443-446): the initialisation upsamples a
     * coarser map (hpm_set_hierarchy_input) and a pixel keeps its new
     * hypothesis only when its cost beats the pre-cost by 0.1
     * (ACMM.cu:1619-1628). */
    int32_t acmm;
    int32_t hierarchy;
} hpm_params;

/* Fills p with the reference defaults. */
static inline void hpm_params_default(hpm_params* p)
{
    p->max_iterations = 3;
    p->patch_size = 11;
    p->num_images = 5;
    p->max_image_size = 3200;
    p->radius_increment = 2;
    p->sigma_spatial = 5.0f;
    p->sigma_color = 3.0f;
    p->top_k = 4;
    p->baseline = 0.54f;
    p->depth_min = 0.0f;
    p->depth_max = 1.0f;
    p->disparity_min = 0.0f;
    p->disparity_max = 1.0f;
    p->geom_consistency = 0;
    p->multi_geometry = 0;
    p->planar_prior = 0;
    p->knn_k = 6;
    p->geom_iterations = 2;
    p->seed_lo = 0x5eed0001u;
    p->seed_hi = 0x00c0ffeeu;
    p->acmm = 0;
    p->hierarchy = 0;
}

#ifdef __cplusplus
}
#endif

#endif /* HPM_TYPES_H_ */
