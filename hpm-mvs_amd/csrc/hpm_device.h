/*
 * hpm_device.h — device-side problem description shared by the HIP kernels
 * (hpm_kernels.hip) and the host runtime (hpm_runtime.cpp).
 *
 * Layout in HBM for one reference view (W x H reference, nsrc sources):
 *   image pool   float, all images back to back, row-major, one grey value
 *                per pixel (the reference converts 8-bit JPEG to float,
 *                HPM.cpp:492-494);
 *   padded pool  per image and texel (x, y), x in [-1, W], y in [-1, H]
 *                (pitch W + 2), the FP16 bilinear coefficients of the 2x2
 *                footprint at (x, y), 8 bytes (hpm_half.h), the border
 *                repeating the edge texels: every clamped source coordinate
 *                in [-1, W] x [-1, H] has its footprint in bounds, so a source
 *                tap is one 8-byte gather with a 32-bit offset, no integer
 *                clamps, and three fused multiply-adds;
 *   depth pool   float, previous-pass depth map of every image (geometric
 *                consistency only, HPM.cpp:561-588);
 *   hyp          float4 per reference pixel: camera-frame plane (n, d)
 *                during an iteration, (world normal, depth) after finalize
 *                (HPM.cu:1751-1764);
 *   cost         float per pixel; sel: uint32 view mask per pixel;
 *   prior_planes float4 per pixel, prior_mask uint32 per pixel (planar prior);
 *   views        hpm_dev_view per image: camera + the camera-only terms of
 *                the plane-induced homography (HPM.cu:229-255).
 */
#ifndef HPM_DEVICE_H_
#define HPM_DEVICE_H_

#include <stdint.h>

#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Plane-induced homography of a source view (HPM.cu:227-287) split as
 * H = A - b q^T with A = Ks Rrel Kr^-1, b = Ks trel (double, per view) and
 * q = Kr^-T n / w (per hypothesis). */
typedef struct hpm_dev_view {
    double A[9];
    double b[3];
    hpm_camera cam;
    int32_t pitch;      /* row pitch of the padded image in entries (width + 2) */
    int64_t img_off;    /* offset of this image in the image pool (elements) */
    int64_t depth_off;  /* offset of this view's depth map in the depth pool */
    int64_t pad_off;    /* offset of entry (-1, -1) of this image in the padded pool */
} hpm_dev_view;

typedef struct hpm_dev_problem {
    double rk_ifx, rk_ify;     /* 1/fx, 1/fy of the reference camera */
    double rk_cx, rk_cy;
    int32_t W, H, nsrc;
    int32_t yy_limit;          /* checkerboard row-pair extent (HPM.cu:1926) */
    hpm_params params;
    uint32_t call_index;       /* RNG stream of this RunPatchMatch call */
    uint32_t salt;             /* RNG stream: reference image id */
    float good[3][3];          /* JudgeExtend thresholds [iter][round] */
    float thr[3];              /* view-selection thresholds per iteration */
    float fallback[3];         /* expf(thr^2 / -0.32) */
    const float* images;
    const uint32_t* images_pad; /* padded pool: entry (x, y) = FP16 coefficients c0 | c1 << 16, c2 | c3 << 16 */
    const float* depths;
    const hpm_dev_view* views;
    float* hyp;                /* float4 per pixel */
    float* cost;
    uint32_t* sel;
    /* photometric cost of each pixel's current hypothesis on every source
     * view, [view][pixel]; NaN in view 0's entry when not known (may be
     * null: no cache) */
    float* vcost;
    const float* prior_planes; /* float4 per pixel */
    const uint32_t* prior_mask;
    unsigned long long* n_eval; /* NCC evaluation counter (may be null) */
    /* ACMM hierarchy (params.hierarchy): the coarser map, float4 per coarse
     * pixel (world normal xyz, depth), its JBU window scale, and the
     * pre-cost per pixel */
    const float* coarse;
    int32_t coarse_w, coarse_h, coarse_scale;
    float* pre_cost;
} hpm_dev_problem;

/* Colour-grid mapping of the checkerboard kernels: a wave (64 lanes) takes
 * every second pixel of HPM_WAVE_ROWS consecutive rows, 64 / HPM_WAVE_ROWS
 * lanes per row; a 256-lane block four such waves stacked. */
#ifndef HPM_WAVE_ROWS
#define HPM_WAVE_ROWS 2
#endif
#define HPM_COLOUR_TILE_X (64 / HPM_WAVE_ROWS)  /* colour-grid columns (pixel pairs) per block */
#define HPM_COLOUR_TILE_Y (4 * HPM_WAVE_ROWS)   /* rows per block */

/* One tile of a problem in the batched kernels: 64x4 pixels of the full
 * grid, or HPM_COLOUR_TILE_X x HPM_COLOUR_TILE_Y of the colour grid. */
typedef struct hpm_tile {
    int32_t problem;  /* index into the problem table */
    int32_t bx, by;   /* tile of the problem's full grid (init, finalize) or colour grid (propagate, filter) */
    int32_t pad_;
} hpm_tile;

enum { HPM_STAGE_INIT = 0, HPM_STAGE_PROPAGATE = 1, HPM_STAGE_FINALIZE = 2, HPM_STAGE_FILTER = 3 };

/* Kernel launchers (hpm_kernels.hip). All launch on `stream` and return a
 * hipError_t value as int. */
int hpm_launch_init(const hpm_dev_problem* pb, void* stream);
int hpm_launch_propagate(const hpm_dev_problem* pb, int iter, int colour, void* stream);
int hpm_launch_finalize(const hpm_dev_problem* pb, void* stream);
int hpm_launch_filter(const hpm_dev_problem* pb, int colour, void* stream);
/* Cost vectors for fixed camera-frame planes at given pixels:
 * out[i*nsrc + v]. */
int hpm_launch_cost_vectors(const hpm_dev_problem* pb, const float* planes, const int32_t* pix, int npix,
                            float* out, void* stream);
/* Joint bilateral upsampling: guide w x h, src_nd float4 at sw x sh
 * (normal xyz, depth w), outputs depth (w x h) and float4 normal. */
int hpm_launch_jbu(const float* guide, int w, int h, const float* src_nd, int sw, int sh, int scale_i,
                   float* out_depth, float* out_normal, void* stream);

/* Copies n finished hypotheses (float4) and costs into a pool slot and the
 * depth (.w) into a depth map. */
int hpm_launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, void* stream);
int hpm_launch_store_state(const float* hyp, const float* cost, int64_t n, float* out_hyp, float* out_cost,
                           float* out_depth, void* stream);
/* JBU of a low-resolution prior followed by the reload into camera-frame
 * prior planes and masks (main.cpp:179-204): mask 1 where the upsampled
 * depth lies in [depth_min, depth_max], else 0. */
int hpm_launch_jbu_prior(const float* guide, int w, int h, const float* src_nd, int sw, int sh, int scale_i,
                         const hpm_camera* ref, float depth_min, float depth_max, float* planes, uint32_t* mask,
                         void* stream);

/* One stage of the batched schedule over a tile table: `grid` resident
 * blocks; gran 0: each takes a contiguous range of the ntiles tiles; gran
 * 1, 2 or 4: they pull work items of gran tiles from per-XCD queues
 * (queue: 8 zeroed counters for this launch). pbs, tiles and queue are
 * device arrays. */
int hpm_launch_batch(int stage, const hpm_dev_problem* pbs, const hpm_tile* tiles, int ntiles, int grid, int iter,
                     int colour, int gran, unsigned int* queue, void* stream);
/* learned prior: network input (x: W*H*32 BF16) of a finished plain pass,
 * then prediction + judge into prior planes / masks (head: W*H*32 FP32) */
int hpm_launch_lprior_features(const hpm_dev_problem* pb, void* x, void* stream);
int hpm_launch_lprior_judge(const hpm_dev_problem* pb, const void* x, const float* head, float* planes,
                            uint32_t* mask, unsigned long long* accepted, int use_conf, void* stream);
/* Fusion (main.cpp:547-604): the consistency tests of reference view ref
 * against its sources, one lane per reference pixel. rec_pix[p * nn + j]
 * receives the source pixel of source j when the pair passes the
 * reprojection, depth and angle tests (-1 otherwise), rec_w the pair's
 * exp(-(err + 200 rdd + 10 angle)), any[p] the bitmask of passing sources.
 * cams and off (pixel offset of each view in depth / hyp) are device
 * arrays over all pool views; hyp holds the world normals (float4). */
typedef struct hpm_fusion_args {
    const hpm_camera* cams;
    const int64_t* off;
    const float* depth;
    const float* hyp;
    int32_t ref;
    int32_t nn;
    int32_t src[32];
    int32_t* rec_pix;
    float* rec_w;
    uint32_t* any;
} hpm_fusion_args;
int hpm_launch_fusion_geometry(const hpm_fusion_args* a, int width, int height, void* stream);
/* p[0..n) = 0 as a kernel (graph-capturable, unlike a memset node) */
int hpm_launch_zero_u64(unsigned long long* p, int n, void* stream);
/* The batched schedule in one persistent cooperative launch (work items of
 * gran 1 or 2 tiles from per-XCD queues, grid barriers between the 10
 * stages; queue: 80 zeroed u32 counters). grid must not exceed the resident
 * blocks hpm_batch_persistent_grid reports. */
int hpm_batch_persistent_grid(int device, int gran, int* grid);
int hpm_launch_batch_persistent(const hpm_dev_problem* pbs, int n, const hpm_tile* tiles, int n_full, int n_colour,
                                int gran, unsigned int* queue, int grid, void* stream);

/* kind 0..3: exp/sin/cos/acos of in[i]; kind 4: Philox blocks */
int hpm_launch_selftest(int kind, const void* in, void* out, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HPM_DEVICE_H_ */
