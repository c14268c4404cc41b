/*
 * hpm_fusion.h — depth-map fusion into a coloured point cloud and its PLY
 * file: RunFusion and StoreColorPlyFileBinaryPointCloud of the reference
 * (HPM-MVS/main.cpp:502-633, HPM.cpp:403-456), on the maps a pipeline left
 * in a pool.
 */
#ifndef HPM_FUSION_H
#define HPM_FUSION_H

#include <stdint.h>

#include "hpm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One fused point: world coordinates, world normal, colour (the
 * reference's PointList, main.h:79-83, with the colour as bytes). */
typedef struct hpm_point {
    float x, y, z;
    float nx, ny, nz;
    uint8_t r, g, b, pad_;
} hpm_point;

/* RunFusion (main.cpp:502-633) over the views of a pool, after
 * hpm_pipeline_run: view i of the fusion is pool view i with sources
 * src_ids[i * max_src + j], j < num_src[i] (pair.txt order), for i <
 * num_views. Depths are the geometric-consistency maps when
 * geom_consistency is nonzero (depths_geom.dmb), the plain-pass maps
 * otherwise (depths.dmb); normals are the last pass's (normals.dmb);
 * colours are the pool's grey images (hpm_fusion_run_rgb takes colour
 * images). The points, in the reference's
 * order (views in turn, pixels row-major), are returned in *points (free
 * with hpm_fusion_free) and their number in *num_points. The consistency
 * tests of every (pixel, source) pair run on the device; the masks, whose
 * updates depend on the points kept before (in raster order, with the
 * reference's per-view used_list), are applied on the host in that order.
 * Returns 0, or a negative code with a message in err. */
int hpm_fusion_run(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
                   int geom_consistency, hpm_point** points, int64_t* num_points, char* err, int err_len);
/* RunFusion of ACMM+NESP (ACMM+NESP/windows/main.cpp:164-300) on the same
 * inputs: a point needs 2 consistent sources, and its position, normal and
 * colour are the means of the reference pixel's and its consistent
 * sources' (the normal as cv::Vec3f /= float computes it, a product with
 * the float reciprocal; the colour truncated as the reference's PLY writer
 * does, ACMM.cpp:396-398). */
int hpm_fusion_run_acmm(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
                        int geom_consistency, hpm_point** points, int64_t* num_points, char* err, int err_len);
/* Either rule (acmm nonzero: ACMM+NESP's) with the point colours from
 * colors[i], width * height RGB byte triples of pool view i at the pool's
 * size (the reference's IMREAD_COLOR image, main.cpp:523, 573); a NULL
 * array or entry takes the grey image. */
int hpm_fusion_run_rgb(hpm_pool* pool, int num_views, const int* num_src, const int* src_ids, int max_src,
                       int geom_consistency, const uint8_t* const* colors, int acmm, hpm_point** points,
                       int64_t* num_points, char* err, int err_len);
void hpm_fusion_free(hpm_point* points);

/* RunFusion as the reference's main calls it (main.cpp:675): reads
 * pair.txt, cams/%08d_cam.txt, images/%08d.jpg or .pgm (grey for the
 * consistency tests' cameras, colour for the points: hpm_read_image_rgb) and
 * HPM/2333_%08d/{depths_geom.dmb or depths.dmb, normals.dmb} of every view
 * (views 0..n-1 in pair.txt order), fuses them with hpm_fusion_run and
 * writes HPM/HPM_model.ply. *num_points (may be NULL) receives the point
 * count. Returns 0, or a negative code with a message in err. */
int hpm_fusion_run_folder(const char* dense_folder, int geom_consistency, int64_t* num_points, char* err,
                          int err_len);

/* RunFusion of ACMM+NESP as its main calls it (ACMM+NESP/windows/
 * main.cpp:164-297, 385): the same inputs from ACMM_NESP/2333_%08d, fused
 * with hpm_fusion_run_acmm, written to ACMM_NESP/ACMM_NESP_model.ply. */
int hpm_fusion_run_folder_acmm(const char* dense_folder, int geom_consistency, int64_t* num_points, char* err,
                               int err_len);

/* StoreColorPlyFileBinaryPointCloud (HPM.cpp:403-456): binary little-endian
 * PLY, vertex x y z nx ny nz (float) red green blue (uchar), non-finite
 * coordinates written as 0; points in array order (the reference writes
 * them in an OpenMP order). Returns 0 or -1. */
int hpm_write_ply(const char* path, const hpm_point* points, int64_t num_points);

#ifdef __cplusplus
}
#endif

#endif
