/*
 * hpm_prior.h — host construction of the planar prior (HPM-MVS/HPM.cpp:
 * 759-1179, main.cpp:73-480) without OpenCV or PCL.
 *
 * Replacements, each deterministic and documented:
 *   cv::Subdiv2D        incremental Delaunay triangulation with exact integer
 *                       predicates (128-bit in-circle test); points on a
 *                       common circle are not inserted into each other's
 *                       cavity, which fixes one triangulation of cocircular
 *                       lattice points;
 *   cv::SVD::solveZ     the 3-point plane as a cross product in FP64;
 *   pcl::KdTreeFLANN    an exact grid search, ties broken by (distance,
 *                       support-point index);
 *   JudgeNotOnline      exact integer collinearity (the reference uses a
 *                       float Heron area, HPM.cpp:928-937).
 * Points are integer pixel coordinates (x, y) as int32 pairs. Planes are
 * camera-frame (n, d) float4 with n.X + d = 0, as the kernels use them.
 */
#ifndef HPM_PRIOR_H_
#define HPM_PRIOR_H_

#include <stdint.h>

#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GetSupportPoints / GetSupportPointsScale0/1 (HPM.cpp:759-785, 1125-1179):
 * per 5x5 block (column-major block and pixel order) the first strict
 * minimum of cost among costs < 2, kept if < 0.1. Returns the number of
 * points; writes at most max_points pairs. */
int hpm_support_points(const float* cost, int width, int height, int32_t* points_xy, int max_points);

/* DelaunayTriangulation (HPM.cpp:787-809): triangles whose three vertices
 * are input points (indices into points_xy, positively oriented:
 * (b - a) x (c - a) > 0). Co-circular ties are broken by point index (a
 * symbolic perturbation of the lifted coordinate), so the triangulation is
 * unique; of points with equal coordinates the first is used. Canonical
 * order: by smallest vertex index, listed first, then around that vertex
 * from its nearest neighbour (or its clockwise hull edge). Coordinate
 * differences must stay below 2^14. Returns the triangle count or a
 * negative error; writes at most max_tris triples. */
int hpm_delaunay(const int32_t* points_xy, int n, int32_t* tris, int max_tris);

/* Triangle rasterisation of main.cpp:96-116: barycentric stepping with
 * step 1/(longest edge), later triangles overwrite; mask = triangle index
 * + 1 (float, as the reference's mask). */
void hpm_rasterize_triangles(const int32_t* points_xy, const int32_t* tris, int ntris, float* mask, int width,
                             int height);

/* GetPriorPlaneParams_factor (HPM.cpp:811-844) with
 * Get3DPointonRefCam_factor (HPM.cpp:236-245): the plane through the three
 * back-projected vertices, normal scaled so that d >= 0. */
void hpm_plane_from_points(const hpm_camera* ref, const int32_t p[6], const float* depth, int width, float factor,
                           float out[4]);

/* GetDepthFromPlaneParam_factor (HPM.cpp:846-849) */
float hpm_depth_from_plane_factor(const hpm_camera* ref, const float plane[4], int x, int y, float factor);

/* DepthsPredictSupplement (HPM.cpp:939-1097): for every pixel whose mask is
 * not > 0, the k nearest support points, the first non-collinear triple in
 * the reference's order, its plane; when the plane's depth at the pixel is
 * within [depth_min, depth_max] the mask becomes -233 and the plane is
 * stored in planes (float4 per pixel). k >= 3; the reference uses 6. */
int hpm_knn_supplement(const hpm_camera* ref, const int32_t* points_xy, int npoints, const float* depth, int width,
                       int height, float factor, int k, float depth_min, float depth_max, float* mask, float* planes,
                       int num_threads);

/* ComputingMultiScaleSettings (HPM.cpp:851-860): halvings of the larger
 * side until it is <= 1000. */
int hpm_multiscale_count(int width, int height);

#ifdef __cplusplus
}
#endif

#endif /* HPM_PRIOR_H_ */
