/*
 * HPM.h — the reference's per-problem C++ class (HPM-MVS/HPM.h:48-112) and
 * its free helpers (HPM.h:6-17), re-exposed over the C-ABI so that a
 * main.cpp-style driver compiles against the same method names. OpenCV and
 * PCL types are replaced by the small value types below (hpm::MatF,
 * hpm::MatV3, hpm::Point, hpm::Rect, hpm::PointCloud); float4 and int2 are
 * plain structs. Method semantics follow the reference, with status codes
 * thrown as hpm::Error instead of exit().
 */
#ifndef HPM_MVS_HPM_H_
#define HPM_MVS_HPM_H_

#include <stdexcept>
#include <string>
#include <vector>

#include "hpm_capi.h"
#include "hpm_io.h"
#include "hpm_prior.h"
#include "hpm_types.h"

namespace hpm {

struct float4 {
    float x, y, z, w;
};
struct int2 {
    int x, y;
};
struct Point {
    int x, y;
};
struct Rect {
    int x, y, width, height;
    bool contains(const Point& p) const { return p.x >= x && p.x < x + width && p.y >= y && p.y < y + height; }
};
struct Triangle {
    Point pt1, pt2, pt3;
    Triangle(Point a, Point b, Point c) : pt1(a), pt2(b), pt3(c) {}
};

/* row-major single- and three-channel float images (cv::Mat_<float>,
 * cv::Mat_<cv::Vec3f>) */
struct MatF {
    int rows = 0, cols = 0;
    std::vector<float> data;
    MatF() = default;
    MatF(int r, int c, float v = 0.0f) : rows(r), cols(c), data((size_t)r * c, v) {}
    float& operator()(int r, int c) { return data[(size_t)r * cols + c]; }
    float operator()(int r, int c) const { return data[(size_t)r * cols + c]; }
};
struct MatV3 {
    int rows = 0, cols = 0;
    std::vector<float> data;
    MatV3() = default;
    MatV3(int r, int c) : rows(r), cols(c), data((size_t)r * c * 3, 0.0f) {}
    float* operator()(int r, int c) { return &data[3 * ((size_t)r * cols + c)]; }
    const float* operator()(int r, int c) const { return &data[3 * ((size_t)r * cols + c)]; }
};
/* pcl::PointCloud<pcl::PointXY> */
struct PointCloud {
    std::vector<Point> points;
    void push_back(Point p) { points.push_back(p); }
    void clear() { points.clear(); }
};
typedef PointCloud* PointCloudPtr;

/* `Problem` (main.h:54-57); the last three fields are ACMM+NESP's
 * (ACMM+NESP/windows/main.h:46-52), used by hpm::ACMM */
struct Problem {
    int ref_image_id = 0;
    std::vector<int> src_image_ids;
    int max_image_size = 3200;
    int num_downscale = 0;
    int cur_image_size = 3200;
};

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

int readDepthDmb(const std::string& file_path, MatF& depth);
int readNormalDmb(const std::string& file_path, MatV3& normal);
int writeDepthDmb(const std::string& file_path, const MatF& depth);
int writeNormalDmb(const std::string& file_path, const MatV3& normal);
hpm_camera ReadCamera(const std::string& cam_path);

class HPM {
public:
    HPM();
    ~HPM();
    HPM(const HPM&) = delete;
    HPM& operator=(const HPM&) = delete;

    void InuputInitialization(const std::string& dense_folder, const Problem& problem);
    void CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem);
    void RunPatchMatch();
    void SetGeomConsistencyParams(bool multi_geometry = false);
    void SetPlanarPriorParams();
    int GetReferenceImageWidth();
    int GetReferenceImageHeight();
    MatF GetReferenceImage();
    float4 GetPlaneHypothesis(const int index);
    float GetCost(const int index);
    void GetSupportPoints(std::vector<Point>& support2DPoints);
    std::vector<Triangle> DelaunayTriangulation(const Rect boundRC, const std::vector<Point>& points);
    float4 GetPriorPlaneParams_factor(const Triangle triangle, const MatF& depths, float factor);
    float GetDepthFromPlaneParam_factor(const float4 plane_hypothesis, const int x, const int y, float factor);
    float GetMinDepth();
    float GetMaxDepth();
    void CudaPlanarPriorInitialization(const std::vector<float4>& PlaneParams, const MatF& masks);
    int ComputingMultiScaleSettings(int width, int height);
    void WritePointCloud(PointCloudPtr& cloud, std::vector<Point> points);
    float4 TransformNormal(float4 plane_hypothesis);
    float4 TransformNormal2RefCam(float4 plane_hypothesis);
    float GetDistance2Origin_factor(const int2 p, const float depth, const float4 normal, float factor);
    void ReloadPlanarPriorInitialization(const MatF& masks, float4* prior_plane_parameters);
    void DepthsPredictSupplement(PointCloudPtr cloud, MatF& depth, MatV3& normal_reliable, float4* plane_parameters,
                                 MatF& mask, float factor);
    void CudaPlanarPriorInitializationSupplement(const std::vector<float4>& PlaneParams, MatF& masks,
                                                 const float4* plane_hypothesis);
    void GetSupportPointsScale0(std::vector<Point>& support2DPoints, MatF costs);
    void GetSupportPointsScale1(std::vector<Point>& support2DPoints, MatF costs);
    void JointBilateralUpsampling(const MatF& scaled_image_float, const MatF& src_depthmap, MatF& upsample_depthmap,
                                  const MatV3& src_normal, MatV3& upsample_normal);
    void ReleasePriorCudaMemory();
    void ReleaseProblemCudaMemory();
    void ReleaseProblemHostMemory();

    /* settings the reference hard-codes (KNN K, seeds) */
    hpm_params& params() { return params_; }
    hpm_handle* handle() { return h_; }

private:
    void check(int rc, const char* what);
    void support_points_of(const MatF& costs, std::vector<Point>& out);
    hpm_handle* h_;
    hpm_params params_;
    hpm_camera ref_cam_;
};

/* `class ACMM` of ACMM+NESP (ACMM+NESP/windows/ACMM.h:47-62) over a handle
 * with the ACMM kernel variant (hpm_params.acmm): the same method names and
 * call order as the reference's ProcessProblem (ACMM+NESP/windows/
 * main.cpp:74-129). Results go to and come from dense_folder/ACMM_NESP.
 * For the hierarchy the coarser level's normals.dmb and depths_geom.dmb are
 * read and upsampled inside RunPatchMatch (the reference reads the JBU
 * output RunJBU wrote to depths.dmb; the two are the same upsampling). */
class ACMM {
public:
    ACMM();
    ~ACMM();
    ACMM(const ACMM&) = delete;
    ACMM& operator=(const ACMM&) = delete;

    void InuputInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, const int idx);
    void CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem);
    void RunPatchMatch();
    void SetGeomConsistencyParams(bool multi_geometry = false);
    void SetHierarchyParams();
    int GetReferenceImageWidth();
    int GetReferenceImageHeight();
    MatF GetReferenceImage();
    float4 GetPlaneHypothesis(const int index);
    float GetCost(const int index);

    hpm_params& params() { return params_; }
    hpm_handle* handle() { return h_; }

private:
    void check(int rc, const char* what);
    hpm_handle* h_;
    hpm_params params_;
    hpm_camera ref_cam_;
    bool hierarchy_ = false;
};

} // namespace hpm

#endif /* HPM_MVS_HPM_H_ */
