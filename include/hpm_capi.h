/*
 * hpm_capi.h — the C-ABI drop-in boundary of the MI355X PatchMatch path.
 *
 * One opaque handle per reference-view problem, mirroring the reference's
 * per-problem class `HPM` (HPM-MVS/HPM.h:48-112). Every entry point returns
 * 0 on success and a negative status otherwise; hpm_last_error() returns the
 * message of the last failure on that handle. Nothing here exits the process
 * (the reference's CudaSafeCall does, HPM.cpp:64-72).
 *
 * Semantics kept from the reference:
 *   - hpm_run_patchmatch is synchronous: on return the host mirror holds the
 *     (world normal, depth) hypotheses and costs (HPM.cu:1954-1956);
 *   - the handle is stateful: a prior or geometric-consistency run starts
 *     from the hypotheses and costs already on the device (HPM.cu:523-552);
 *   - a handle is not thread-safe; all its work runs on one HIP stream.
 *
 * Arrays are plain host pointers: float4 values are 4 consecutive floats.
 */
#ifndef HPM_CAPI_H_
#define HPM_CAPI_H_

#include <stdint.h>

#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpm_handle hpm_handle;
typedef struct hpm_pool hpm_pool;

#define HPM_OK 0
#define HPM_EINVAL (-1)   /* bad argument or call order */
#define HPM_EIO (-2)      /* file missing or malformed */
#define HPM_EDEVICE (-3)  /* HIP runtime error */
#define HPM_ENOMEM (-4)   /* allocation failure */
#define HPM_ETOOBIG (-5)  /* input larger than a fixed-size entry point admits */

/* Timing and work counters of the last hpm_run_patchmatch. */
typedef struct hpm_run_stats {
    uint64_t n_eval;          /* (pixel, hypothesis, source view) NCC evaluations */
    uint64_t n_eval_propagate; /* the part evaluated by the 6 propagation launches */
    float ms_total;           /* init .. filter, device time (hipEvents) */
    float ms_init;
    float ms_propagate;       /* all 6 red/black propagation launches */
    float ms_finalize_filter;
    int32_t launches;
} hpm_run_stats;

/* Every view of a scene resident on one device: images uploaded once
 * (downscaled above max_image_size with the camera, HPM.cpp:524-551), plus
 * per view the last plain-pass depth map (which 0, the reference's
 * depths.dmb), the last geometric-consistency depth map (which 1,
 * depths_geom.dmb) and the last hypotheses and costs. No reference
 * counterpart: the reference re-reads images and .dmb files per problem.
 * Returns NULL on bad arguments, device errors, or when the padded source
 * copies, (width + 3) * (height + 2) entries per view, exceed 2^29 entries
 * together (the kernels address them with 31-bit byte offsets; split larger
 * scenes into several pools) or a view's width or height exceeds 2^21 - 3
 * or 2^23 - 3 pixels. Source texels are sampled as their nearest FP16
 * values (exact for 8-bit grey images). */
hpm_pool* hpm_pool_create(int device, int num_views, const hpm_camera* cams, const float* const* images,
                          int max_image_size);
void hpm_pool_destroy(hpm_pool* pool);
int hpm_pool_camera(const hpm_pool* pool, int view, hpm_camera* out);
int hpm_pool_set_depth(hpm_pool* pool, int view, int which, const float* depth);
int hpm_pool_get_depth(hpm_pool* pool, int view, int which, float* depth);

/* HPM::HPM / HPM::~HPM (HPM.cpp:99-133). device < 0 selects device 0. */
hpm_handle* hpm_create(int device);
void hpm_destroy(hpm_handle* h);
const char* hpm_last_error(const hpm_handle* h);

/* PatchMatchParams (HPM.h:29-46) plus the settings the reference hard-codes
 * (hpm_types.h). hpm_get_params after input init returns the relaxed depth
 * range. */
int hpm_set_params(hpm_handle* h, const hpm_params* p);
int hpm_get_params(const hpm_handle* h, hpm_params* p);

/* HPM::SetGeomConsistencyParams (HPM.cpp:468-475) */
int hpm_set_geom_consistency(hpm_handle* h, int multi_geometry);
/* HPM::SetPlanarPriorParams (HPM.cpp:477-480) */
int hpm_set_planar_prior(hpm_handle* h);
/* ACMM::SetHierarchyParams (ACMM+NESP/windows/ACMM.cpp:443-446) with the
 * hierarchy branch of ACMM::CudaSpaceInitialization (ACMM.cpp:660-717), after
 * hpm_device_init, on a handle with params.acmm set: coarse_world4 holds
 * (world normal xyz, depth) per pixel of the previous, coarser scale
 * (normals.dmb and depths_geom.dmb there). RunPatchMatch then starts from
 * their joint bilateral upsampling guided by the reference image (RunJBU,
 * ACMM.cpp:804-853, and the upsample branch of RandomInitialization,
 * ACMM.cu:532-601) and keeps a pixel's new hypothesis only where its cost
 * beats the pre-cost by 0.1 (ACMM.cu:1619-1628). The coarse map must be
 * smaller by an integer factor of 2 or more. */
int hpm_set_hierarchy_input(hpm_handle* h, const float* coarse_world4, int coarse_w, int coarse_h);

/* HPM::InuputInitialization (HPM.cpp:482-589): reads images/%08d.{pgm,jpg}
 * as grey, cams/%08d_cam.txt, downscales images above max_image_size, sets
 * the relaxed depth range; with geometric consistency also reads
 * HPM/2333_%08d/depths.dmb (or depths_geom.dmb) for every view. Images are
 * images/%08d.jpg (sequential or progressive Huffman JPEG, decoded to the
 * luminance that cv::imread(IMREAD_GRAYSCALE) returns), else
 * images/%08d.pgm (8-bit P5). */
int hpm_input_init(hpm_handle* h, const char* dense_folder, int ref_image_id, const int* src_image_ids,
                   int num_src);
/* The same on a pool: the problem's reference and source views are pool
 * views, nothing is copied. The pool must outlive the handle's use of it. */
int hpm_input_init_pool(hpm_handle* h, hpm_pool* pool, int ref_view, const int* src_views, int num_src);
/* In-memory form of the same: images[i] is cams[i].width x cams[i].height
 * row-major grey; depths (may be NULL) gives the previous-pass depth map of
 * every image for geometric consistency. ref_image_id only salts the RNG. */
int hpm_input_init_memory(hpm_handle* h, int ref_image_id, int num_images, const hpm_camera* cams,
                          const float* const* images, const float* const* depths);

/* HPM::CudaSpaceInitialization (HPM.cpp:591-700): device buffers and image
 * upload; with geometric consistency reloads the reference view's previous
 * depths.dmb|depths_geom.dmb, normals.dmb and costs.dmb from dense_folder. */
int hpm_device_init(hpm_handle* h, const char* dense_folder);
/* In-memory form of the geometric-consistency reload: hyp_world holds
 * (world normal xyz, depth) per pixel, cost the previous costs. */
int hpm_device_init_state(hpm_handle* h, const float* hyp_world, const float* cost);

/* Philox stream index of the next RunPatchMatch (each call then advances
 * it by one; hpm_device_init resets it to 0). No counterpart in the
 * reference, which seeds every launch from clock64() (HPM.cu:521): it lets
 * a caller that drives the class API reproduce a schedule's passes bit for
 * bit (hpm_pipeline_options.stream_base). Call after hpm_device_init. */
int hpm_set_stream_index(hpm_handle* h, uint32_t index);

/* HPM::RunPatchMatch (HPM.cu:1906-1957) */
int hpm_run_patchmatch(hpm_handle* h);
int hpm_get_run_stats(const hpm_handle* h, hpm_run_stats* out);

/* HPM::GetReferenceImageWidth / Height (HPM.cpp:724-732) */
int hpm_dims(const hpm_handle* h, int* width, int* height);
/* Camera of image i of the problem (0 = reference) after input init,
 * including the max_image_size downscale (HPM.cpp:524-551). */
int hpm_get_camera(const hpm_handle* h, int image, hpm_camera* out);
/* HPM::GetReferenceImage (HPM.cpp:734-737): copies W*H floats. */
int hpm_copy_reference_image(const hpm_handle* h, float* out);
/* HPM::GetPlaneHypothesis / GetCost (HPM.cpp:739-747), one pixel. */
int hpm_get_plane(const hpm_handle* h, int index, float out[4]);
float hpm_get_cost(const hpm_handle* h, int index);
/* Whole-map copies of the host mirror (the readback loops of
 * main.cpp:63-71). */
int hpm_copy_hypotheses(const hpm_handle* h, float* out4);
int hpm_copy_costs(const hpm_handle* h, float* out);
/* HPM::GetMinDepth / GetMaxDepth (HPM.cpp:749-757) */
int hpm_depth_range(const hpm_handle* h, float* depth_min, float* depth_max);

/* HPM::ReloadPlanarPriorInitialization (HPM.cpp:901-921) and
 * CudaPlanarPriorInitialization (HPM.cpp:702-722): per-pixel camera-frame
 * prior planes (n, d) and masks; planes are read only where mask > 0. */
int hpm_upload_prior(hpm_handle* h, const float* planes4, const uint32_t* mask);
/* One scale of the planar prior built on the device from the handle's last
 * RunPatchMatch, as main.cpp builds it on the host (main.cpp:84-162 at
 * factor 0.25, 238-314 at 0.5, 387-454 at 1; HPM.cpp:759-1123), every
 * stage on the device: support points of the (resized) cost map, their
 * Delaunay triangulation (the one of hpm_delaunay), the rasterised triangle
 * mask, the triangle planes, the KNN supplement with knn_k neighbours, then
 * at factor 1 the range-gated planes and masks, below it their world
 * normals and depths upsampled by JBU (JointBilateralUpsampling,
 * HPM.cpp:1215-1281) into the prior. The prior is left on the handle with
 * the planar prior enabled, equal bit for bit to the host construction.
 * ms_host (may be NULL) accumulates the milliseconds the host waits for
 * the point and triangle counts, which size the later launches. */
int hpm_build_planar_prior(hpm_handle* h, float factor, int knn_k, double* ms_host);
/* HPM::ReleasePriorCudaMemory (HPM.cpp:1283-1286); idempotent. */
int hpm_release_prior(hpm_handle* h);
/* HPM::ReleaseProblemCudaMemory / ReleaseProblemHostMemory
 * (HPM.cpp:1288-1307); idempotent. */
int hpm_release(hpm_handle* h);

/* HPM::JointBilateralUpsampling (HPM.cpp:1215-1281) + JBU_cu
 * (HPM.cu:1984-2052): guide w x h grey image, src_nd float4 (world normal
 * xyz, depth) at sw x sh. Writes out_depth (w x h) and out_normal (float4,
 * w = 0). Returns 1 and writes nothing when the integer scale is 1. Does
 * not need input init. */
int hpm_jbu(hpm_handle* h, const float* guide, int w, int hgt, const float* src_nd, int sw, int sh,
            float* out_depth, float* out_normal);

/* Multi-view NCC cost vectors of fixed camera-frame planes at pixels
 * (ComputeMultiViewCostVector, HPM.cu:438-443): out[i*nsrc + v]. Needs
 * device init. */
int hpm_cost_vectors(hpm_handle* h, const float* planes4, const int32_t* pix_xy, int npix, float* out);

/* Stages of hpm_run_patchmatch for step-by-step comparison:
 * stage 0 init, 1 propagate(iter, colour), 2 finalize, 3 filter(colour),
 * 4 copy the device state to the host mirror. */
int hpm_debug_stage(hpm_handle* h, int stage, int iter, int colour);
/* Device state without finalize: hyp (4 floats/px), cost, view masks. */
int hpm_debug_read_state(hpm_handle* h, float* hyp4, float* cost, uint32_t* selmask);
int hpm_debug_write_state(hpm_handle* h, const float* hyp4, const float* cost, const uint32_t* selmask);
/* The device prior of the handle: camera-frame planes (n, d; 4 floats/px)
 * and masks, as a prior run reads them. Fails when no prior is on the
 * device. */
int hpm_debug_read_prior(hpm_handle* h, float* planes4, uint32_t* mask);

/* Evaluates the shared device arithmetic for comparison with a host build:
 * kind 0..3 = exp, sin, cos, acos of the n floats in `in` (out: n floats);
 * kind 4 = Philox4x32-10 of n blocks, each 6 uint32 (ctr0..3, key0..1) in
 * `in`, 4 uint32 out per block; kind 5, 6 = the tap fraction and floor of
 * n floats; kind 7 = the NCC's square root of n floats; kind 8 = the NCC's
 * quotient of n (numerator, denominator) float pairs (out: n floats);
 * kind 9 = the warp's double reciprocal of n doubles (out: n doubles). No
 * reference counterpart: a check of the replacement for cuRAND and
 * fast-math intrinsics and of the correctly rounded device operations. */
int hpm_device_selftest(hpm_handle* h, int kind, const void* in, void* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* HPM_CAPI_H_ */
