/*
 * hpm_internal.h — the runtime objects behind the C-ABI handles, shared by
 * hpm_runtime.cpp (the per-problem `HPM` equivalent) and hpm_pipeline.cpp
 * (the `ProcessProblem` / `main` pass schedule).
 *
 * hpm_pool: every view of a scene resident on one device. Images are
 * uploaded once; per view the pool also keeps the last plain-pass depth map
 * (the reference's depths.dmb), the last geometric-consistency depth map
 * (depths_geom.dmb) and the last (world normal, depth) hypotheses and costs
 * (normals.dmb, costs.dmb), so passes hand results to each other on the
 * device instead of through files (HPM.cpp:561-588, 662-699).
 */
#ifndef HPM_INTERNAL_H_
#define HPM_INTERNAL_H_

#include <hip/hip_runtime_api.h>

#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "hpm_capi.h"
#include "hpm_device.h"

struct hpm_pool {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0;
    std::vector<hpm_camera> cams;             /* after the max_image_size downscale */
    std::vector<std::vector<float>> images;   /* host copies (JBU guide, GetReferenceImage) */
    std::vector<int64_t> off;                 /* pixel offset of each view in the device arrays */
    int64_t total = 0;
    float* d_images = nullptr;
    std::vector<int64_t> pad_off;             /* offset of texel (-1, -1) of each view, padded pool */
    int64_t pad_total = 0;
    uint32_t* d_images_pad = nullptr;         /* entry (x, y), 8 bytes: the FP16 bilinear coefficients of
                                                 texels (x..x+1, y..y+1) (hpm_half.h); x in [-1, W],
                                                 y in [-1, H], edges replicated */
    float* d_depth_plain = nullptr;           /* depths.dmb of each view */
    float* d_depth_geom = nullptr;            /* depths_geom.dmb of each view */
    float* d_state_hyp = nullptr;             /* float4: world normal, depth of the last pass */
    float* d_state_cost = nullptr;
    std::vector<char> have_state;
    std::string err;
    /* problem handles of finished pass schedules over this pool, with their
     * streams and device buffers, reused by the next schedule (creating a
     * stream and allocating per problem cost more than a 1600x1066 launch);
     * and the learned prior the last schedule loaded, keyed by its file.
     * Schedules over one pool do not run concurrently: they write its maps. */
    std::mutex spare_mu;
    std::vector<hpm_handle*> spare;
    std::string lprior_key;
    struct hpm_lprior* lprior = nullptr;
};

struct hpm_handle {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hpm_params params;
    int ref_id = 0;
    int num_images = 0;

    hpm_pool* pool = nullptr;   /* images of this problem */
    bool own_pool = false;
    std::vector<int> views;     /* pool view of image i (0 = reference) */
    std::vector<hpm_camera> cams;

    bool have_input = false;
    bool have_device = false;
    uint32_t call_index = 0;

    /* per-problem device state */
    hpm_dev_view* d_views = nullptr;
    float* d_hyp = nullptr;
    float* d_cost = nullptr;
    uint32_t* d_sel = nullptr;
    float* d_vcost = nullptr;     /* cost cache, nsrc x pixels (hpm_dev_problem::vcost) */
    size_t vcost_cap = 0;         /* floats */
    float* d_prior = nullptr;
    uint32_t* d_mask = nullptr;
    unsigned long long* d_neval = nullptr;
    float* d_jbu_src = nullptr;   /* low-resolution prior input of the JBU */
    size_t jbu_src_cap = 0;       /* bytes */
    float* d_coarse = nullptr;    /* ACMM hierarchy: coarser map, float4 (world normal, depth) */
    float* d_pre_cost = nullptr;  /* ACMM hierarchy: pre-cost per pixel */
    size_t coarse_cap = 0;        /* coarse pixels */
    int coarse_w = 0, coarse_h = 0, coarse_scale = 0;
    bool have_coarse = false;
    void* d_ws = nullptr;         /* learned-prior activations */
    size_t ws_cap = 0;            /* bytes */
    size_t alloc_pixels = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool have_prior = false;

    /* host mirrors, sized and filled by the first readback after a device
     * initialisation; until then the getters read zeros */
    std::vector<float> h_hyp;
    std::vector<float> h_cost;
    bool mirror_filled = false;
    hpm_run_stats stats;
};

/* error helpers */
int hpm_fail(hpm_handle* h, int code, const char* fmt, ...);

/* returns HPM_EDEVICE with the HIP error in the handle's message */
#define HPM_HIP(h, call)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            return hpm_fail((h), HPM_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                                    \
        }                                                                                           \
    } while (0)

/* The tiles of the batched and work-queue schedules: 64x4 tiles of every
 * problem's full grid (init, finalize) and colour-grid tiles (propagate,
 * filter), problem by problem, rows outer (hpm_batch.cpp). */
void hpm_make_tiles(const hpm_dev_problem* pbs, int n, std::vector<hpm_tile>* full, std::vector<hpm_tile>* colour);

/* The device problem description of a ready handle (its pool, views, state,
 * call index), and the readiness check of RunPatchMatch. */
void hpm_fill_problem(const hpm_handle* h, hpm_dev_problem* pb);
int hpm_ensure_ready(hpm_handle* h);
/* Marks the handle's cost cache unknown for every pixel: required after
 * anything but the init and propagation kernels writes d_hyp. */
int hpm_invalidate_cost_cache(hpm_handle* h);
/* Allocates the prior planes / masks if needed and enables the planar prior
 * as if a prior had been uploaded (the caller fills d_prior / d_mask). */
int hpm_prior_on_device(hpm_handle* h);
/* A device scratch area of at least `bytes`, kept with the handle. */
int hpm_workspace(hpm_handle* h, size_t bytes, void** ws);

/* RunPatchMatch with an optional host readback (the C-ABI always reads
 * back; the pass schedule skips it where the host does not need the maps). */
int hpm_run_internal(hpm_handle* h, bool readback);
/* hpm_set_hierarchy_input from a host or a device buffer. */
int hpm_set_hierarchy_internal(hpm_handle* h, const float* coarse_world4, int coarse_w, int coarse_h,
                               bool on_device);
/* Copies the handle's finished (world normal, depth) hypotheses and costs
 * into the pool slot of its reference view and the depth into the plain
 * (which_depth 0) or geometric (1) depth map of that view. */
int hpm_store_to_pool(hpm_handle* h, int which_depth);
/* Loads the pool's last state of the reference view as the starting state
 * of a geometric-consistency run (HPM.cpp:662-699). */
int hpm_load_from_pool(hpm_handle* h);
/* JBU of a low-resolution prior (float4 world normal + depth, sw x sh, on
 * the host) guided by the reference image already on the device, followed by
 * the reload of main.cpp:179-204 on the device: mask 1 where the upsampled
 * depth is within the relaxed range, camera-frame planes there. Leaves the
 * prior uploaded on the handle. Returns 1 (nothing done) when the integer
 * scale is 1. */
int hpm_jbu_prior(hpm_handle* h, const float* src_nd, int sw, int sh, float factor);

/* The learned prior's network over a caller-provided activation area
 * (hpm_lprior.cpp): bytes needed at width x height (input and head not
 * included), the device the weights live on, and the forward pass into
 * `ws` (asynchronous on `stream`). */
struct hpm_lprior;
size_t hpm_lprior_activation_bytes(int width, int height);
int hpm_lprior_device(const hpm_lprior* lp);
hipError_t hpm_lprior_forward_ws(const hpm_lprior* lp, const void* x, int height, int width, float* head, void* ws,
                                 hipStream_t stream);

#endif /* HPM_INTERNAL_H_ */
