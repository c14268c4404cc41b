/*
 * hpm_oracle.h — C interface of the CPU restatement of HPM-MVS's
 * per-reference-view PatchMatch (HPM-MVS/HPM.cu:1-2052).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is
 * compared against; nothing in the product links or calls it. It is used by
 * tests/, by the smoke check and by the CPU-baseline leg of bench.py.
 *
 * The restatement is plain C++ with OpenMP over pixels. It follows the
 * reference's text (function by function, cited in hpm_oracle.cpp) with the
 * same floating-point expression order as the HIP kernels, so that for the
 * same inputs, seed and call index the two produce the same bits wherever
 * the arithmetic is IEEE-exact (all of it: +, -, *, /, sqrt, explicit fma,
 * and the shared elementary functions of hpm_fmath.h). An FP64 variant of
 * the bilateral NCC cost (hpm_oracle_cost_vectors with fp64=1) evaluates the
 * reference formula E[xy]-E[x]E[y] in double precision as the accuracy
 * anchor for the FP32 cost.
 */
#ifndef HPM_ORACLE_H_
#define HPM_ORACLE_H_

#include <stdint.h>

#include "hpm_fusion.h"
#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Micro-scene hooks (tests only). A cost hook, when set, replaces the
 * bilateral NCC (HPM.cu:325-397) of one (reference pixel, camera-frame
 * plane (n, d), source view 1..N-1) everywhere the restatement evaluates
 * it, so that a test can build the cost fields the sampling, view
 * selection, adoption and refinement logic acts on. Call with one thread
 * when the hook is not thread-safe. */
typedef float (*hpm_oracle_cost_fn)(void* user, int32_t px, int32_t py, int32_t view, const float* plane);

/* What one pixel's CheckerboardPropagation + PlaneHypothesisRefinement
 * (HPM.cu:1222-1725, 1143-1220) decided, recorded when trace->px/py is
 * updated by a half-iteration. Directions in the reference's order:
 * 0 left_up, 1 up_far, 2 right_up, 3 down_far, 4 right_down, 5 left_far,
 * 6 left_down, 7 right_far. */
typedef struct hpm_oracle_trace {
    int32_t px, py;           /* in: the pixel to record */
    int32_t hit;              /* out: 1 once recorded */
    int32_t flag[8];          /* direction valid */
    int32_t pos_scan[8];      /* neighbour pixel index chosen by each scan */
    int32_t judge[3];         /* JudgeExtend per round: 1 extend, 0 stop, -1 not reached */
    int32_t pos[8];           /* positions after the extension rounds */
    float ca[8][32];          /* cost_array after the extension rounds */
    float vprior[32];         /* view_selection_priors */
    float pdf[32];            /* sampling_probs (times the prior), before the CDF */
    float cdf[32];            /* after TransformPDFToCDF */
    float draws[15];          /* curand_uniform - FLT_EPSILON */
    float vw[32];             /* view_weights */
    uint32_t tsel;            /* temp_selected_views */
    float wn;                 /* weight_norm */
    float fc[8];              /* final_costs */
    int32_t min_idx;          /* FindMinCostIndex(final_costs) */
    float cost_now;           /* re-scored current hypothesis */
    float rfc[8];             /* restricted_final_costs (prior mask > 0) */
    int32_t max_idx;          /* FindMaxCostIndex(restricted_final_costs), -1 without */
    float rcn;                /* restricted_cost_now */
    int32_t adopted;          /* direction whose hypothesis was adopted, -1 none */
    int32_t nan_fallback;     /* 1 when the NaN-cost fallback adopted */
    float restricted_cost;    /* restricted_cost entering the refinement */
    float cand[5][4];         /* refinement planes (camera frame, d from GetDistance2Origin) */
    float cand_depth[5];      /* depths[i] of the refinement */
    float cand_cost[5];       /* temp_cost */
    float cand_rcost[5];      /* restricted_temp_cost (prior), else 0 */
    int32_t cand_accepted[5];
} hpm_oracle_trace;

typedef struct hpm_oracle_problem {
    int32_t num_images;                      /* 1 reference + sources */
    hpm_camera cams[HPM_MAX_IMAGES];         /* width/height filled */
    const float* images[HPM_MAX_IMAGES];     /* row-major grey, cams[i].width x cams[i].height */
    const float* depths[HPM_MAX_IMAGES];     /* geometric consistency: previous depth maps (NULL otherwise) */
    hpm_params params;                       /* depth_min/depth_max already relaxed */
    const float* prior_planes;               /* planar prior: float4 per ref pixel (camera-frame n, d) */
    const uint32_t* prior_mask;              /* planar prior: u32 per ref pixel */
    uint32_t call_index;                     /* RunPatchMatch call number on this problem (RNG stream) */
    uint32_t salt;                           /* reference view id (RNG stream) */
    /* ACMM hierarchy (params.hierarchy): the coarser map, float4 per coarse
     * pixel (world normal xyz, depth), and the pre-cost per pixel (written
     * by the initialisation, read by the propagation) */
    const float* coarse;
    int32_t coarse_w, coarse_h;
    float* pre_cost;
    /* micro-scene hooks (NULL in every ordinary run) */
    hpm_oracle_cost_fn cost_hook;
    void* cost_hook_user;
    hpm_oracle_trace* trace;
} hpm_oracle_problem;

/* One full RunPatchMatch (HPM.cu:1906-1957): init, 3 red/black iterations,
 * plane->depth/world normal, red/black median filter. hyp (float4 per pixel),
 * cost and selmask are in/out state, as on the device. n_eval (may be NULL)
 * receives the number of (pixel, hypothesis, source view) NCC evaluations.
 * Returns 0 on success. */
int hpm_oracle_run_patchmatch(const hpm_oracle_problem* pb, float* hyp, float* cost,
                              uint32_t* selmask, uint64_t* n_eval, int num_threads);

/* The stages of hpm_oracle_run_patchmatch, for step-by-step parity. */
int hpm_oracle_init(const hpm_oracle_problem* pb, float* hyp, float* cost, uint32_t* selmask,
                    uint64_t* n_eval, int num_threads);
int hpm_oracle_half_iteration(const hpm_oracle_problem* pb, int iter, int colour, float* hyp,
                              float* cost, uint32_t* selmask, uint64_t* n_eval, int num_threads);
/* The same half-iteration restricted to the pixels of rows [y0, y1): the
 * pixels of one colour read only the other colour's state, so a band
 * gives the whole half-iteration's bits inside it (large views). */
int hpm_oracle_half_iteration_rows(const hpm_oracle_problem* pb, int iter, int colour, int y0, int y1,
                                   float* hyp, float* cost, uint32_t* selmask, uint64_t* n_eval, int num_threads);
int hpm_oracle_finalize(const hpm_oracle_problem* pb, float* hyp, int num_threads);
int hpm_oracle_filter(const hpm_oracle_problem* pb, int colour, float* hyp, const float* cost,
                      int num_threads);

/* Multi-view NCC cost vectors for fixed hypotheses (HPM.cu:325-397,438-443).
 * planes: npix camera-frame planes (n, d); pix: npix (x, y) pairs.
 * out: npix x (num_images-1) costs. fp64: 0 the FP32 arithmetic of the HIP
 * path; 1 the double-precision evaluation of the reference formula; for
 * accuracy evidence only, 2 the reference's own FP32 order (homography in
 * float per call, a float division per tap, raw sums, HPM.cu:227-397) and
 * 3 the HIP path's formulation with its warp in FP32 instead of FP64. */
int hpm_oracle_cost_vectors(const hpm_oracle_problem* pb, const float* planes, const int32_t* pix,
                            int npix, int fp64, double* out, int num_threads);

/* ComputeGeomConsistencyCost (HPM.cu:483-508; capped at 3 instead of 5 with
 * params.acmm, ACMM.cu:494-519) of npix camera-frame planes at pixels pix
 * ((x, y) pairs) against source view v; needs depths[v]. */
int hpm_oracle_geom_cost(const hpm_oracle_problem* pb, int v, const float* planes, const int32_t* pix, int npix,
                         float* out);

/* Joint bilateral upsampling (HPM.cu:1984-2052) of a low-resolution
 * depth/normal map guided by a full-resolution grey image. src_nd is float4
 * (world normal xyz, depth) at sw x sh; out_depth / out_normal (float4, w=0)
 * are at w x h. Returns 1 without writing when the integer scale is 1
 * (HPM.cpp:1218-1222). */
int hpm_oracle_jbu(const float* guide, int w, int h, const float* src_nd, int sw, int sh,
                   float* out_depth, float* out_normal, int num_threads);

/* JBU of a low-resolution prior followed by the reload of main.cpp:179-204
 * (and 330-353): mask 1 where the upsampled depth lies within
 * [depth_min, depth_max], camera-frame plane (R n, -n.X) there
 * (TransformNormal2RefCam, GetDistance2Origin_factor with the
 * full-resolution K, HPM.cpp:882-899); planes 0 and mask 0 elsewhere.
 * Returns 1 without writing when the integer scale is 1. */
int hpm_oracle_jbu_prior(const float* guide, int w, int h, const float* src_nd, int sw, int sh,
                         const hpm_camera* ref, float depth_min, float depth_max, float* planes, uint32_t* mask,
                         int num_threads);

/* Host evaluation of the shared elementary functions (hpm_fmath.h):
 * fn 0 exp, 1 sin, 2 cos, 3 acos; and of the tap sampling's integer and
 * fraction split: fn 5 the bilinear fraction (x - floor(x), capped at the
 * largest float below 1), fn 6 floor(x) as an int32 stored in y's bits. */
int hpm_oracle_fmath(int fn, const float* x, float* y, int n);
/* The value a source texel is sampled as (hpm_half.h): nearest FP16. */
int hpm_oracle_round_f16(const float* x, float* y, int n);
/* Host evaluation of Philox4x32-10 and of the per-pixel draw stream
 * (hpm_philox.h): n uniforms of stream (seed, pixel, stream, salt). */
int hpm_oracle_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int hpm_oracle_uniforms(uint32_t seed_lo, uint32_t seed_hi, uint32_t pixel, uint32_t stream, uint32_t salt,
                        float* out, int n);

/* RunFusion (main.cpp:502-633) on the CPU, sequential in the reference's
 * order: views in turn, pixels row-major, masks set as points are kept,
 * used_list kept per view (main.cpp:556). depths[i] and normals[i] (world,
 * 3 floats/px) and colors[i] (r g b bytes/px) of view i at cams[i]'s size;
 * sources src_ids[i * max_src + j], j < num_src[i]. Writes up to cap points
 * to out and returns how many the fusion kept. */
int64_t hpm_oracle_fusion(int num_views, const hpm_camera* cams, const float* const* depths,
                          const float* const* normals, const uint8_t* const* colors, const int* num_src,
                          const int* src_ids, int max_src, hpm_point* out, int64_t cap);
/* The same with ACMM+NESP's rule (ACMM+NESP/windows/main.cpp:164-300). */
int64_t hpm_oracle_fusion_acmm(int num_views, const hpm_camera* cams, const float* const* depths,
                               const float* const* normals, const uint8_t* const* colors, const int* num_src,
                               const int* src_ids, int max_src, hpm_point* out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* HPM_ORACLE_H_ */
