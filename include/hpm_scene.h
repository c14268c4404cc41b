/*
 * hpm_scene.h — seeded synthetic multi-view scenes with ground truth, in the
 * reference's folder layout (images/, cams/, pair.txt). The reference ships
 * no data; these scenes stand in for the ETH3D / Tanks & Temples folders it
 * was run on (HPM-MVS/README.md).
 *
 * Scene: pinhole cameras on a circular arc of radius `look_depth` centred on
 * the look-at point (0, 0, look_depth); view i sits at angle
 * theta_i = (i - (n-1)/2) * arc_length / (look_depth * (n-1)) about the
 * world y axis and looks at the centre, so the middle view is the world
 * frame. Geometry: a fronto-parallel background plane at depth_max, textured
 * rectangular plane patches and y-rotated boxes, all clipped to the slab
 * depth_min <= z <= depth_max of the world frame. Texture: bilinear value
 * noise on each surface with a lattice pitch of pitch_px pixels at the
 * surface's middle-view depth, mean 128, amplitude amp_high, or amp_low for
 * the objects picked (in seeded order, by middle-view coverage) to make up
 * low_texture_fraction of the middle view. Each image adds i.i.d.
 * N(0, noise_sigma^2) noise and is rounded and clamped to 8 bits.
 */
#ifndef HPM_SCENE_H_
#define HPM_SCENE_H_

#include <stdint.h>

#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpm_scene_config {
    uint32_t seed;
    int32_t width, height;
    float fx, fy, cx, cy;
    int32_t num_views;
    float arc_length;           /* metres along the arc, first to last view */
    float look_depth;           /* arc radius = distance to the look-at point */
    int32_t num_planes;         /* plane patches besides the background */
    int32_t num_boxes;
    float low_texture_fraction; /* target share of middle-view pixels */
    float amp_high, amp_low;    /* texture amplitudes (grey levels) */
    float noise_sigma;          /* per-image Gaussian noise */
    float pitch_px;             /* value-noise lattice pitch in pixels */
    float depth_min, depth_max; /* also written to the camera files */
} hpm_scene_config;

/* Defaults: 1 reference + 4 sources at 640x480 (fx=fy=500), 0.3 m arc,
 * radius 5 m, 4 planes, no boxes, amplitude 60, noise sigma 2, pitch 8 px,
 * depths 2..8 m. */
void hpm_scene_default(hpm_scene_config* cfg);

/* Renders every view. cams: num_views cameras (width/height set).
 * images: num_views * W * H bytes. gt_depth (num_views * W * H, camera z)
 * and gt_normal (num_views * W * H * 3, world frame, facing the camera) may
 * be NULL. low_texture (W*H bytes, may be NULL) marks middle-view pixels on
 * low-texture surfaces. Returns 0 on success. */
int hpm_scene_render(const hpm_scene_config* cfg, hpm_camera* cams, uint8_t* images, float* gt_depth,
                     float* gt_normal, uint8_t* low_texture);

/* Renders and writes dense_folder/{images/%08d.pgm, cams/%08d_cam.txt,
 * pair.txt, gt/%08d_depth.dmb, gt/%08d_normal.dmb}. Every view is a
 * reference with all other views as sources, nearest first. */
int hpm_scene_write_folder(const hpm_scene_config* cfg, const char* dense_folder);

#ifdef __cplusplus
}
#endif

#endif /* HPM_SCENE_H_ */
