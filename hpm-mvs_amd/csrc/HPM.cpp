/*
 * HPM.cpp — the reference's `class HPM` (HPM-MVS/HPM.h:48-112,
 * HPM.cpp:99-1350) over the C-ABI of libhpm_mvs.
 */
#include "HPM.h"

#include <math.h>
#include <string.h>

#include <algorithm>

namespace hpm {

int readDepthDmb(const std::string& file_path, MatF& depth)
{
    int w = 0, h = 0, nb = 0;
    if (hpm_read_dmb(file_path.c_str(), nullptr, &w, &h, &nb) != 0 || nb != 1) {
        return -1;
    }
    depth = MatF(h, w);
    return hpm_read_dmb(file_path.c_str(), depth.data.data(), &w, &h, &nb) == 0 ? 0 : -1;
}

int readNormalDmb(const std::string& file_path, MatV3& normal)
{
    int w = 0, h = 0, nb = 0;
    if (hpm_read_dmb(file_path.c_str(), nullptr, &w, &h, &nb) != 0 || nb != 3) {
        return -1;
    }
    normal = MatV3(h, w);
    return hpm_read_dmb(file_path.c_str(), normal.data.data(), &w, &h, &nb) == 0 ? 0 : -1;
}

int writeDepthDmb(const std::string& file_path, const MatF& depth)
{
    return hpm_write_dmb(file_path.c_str(), depth.data.data(), depth.cols, depth.rows, 1);
}

int writeNormalDmb(const std::string& file_path, const MatV3& normal)
{
    return hpm_write_dmb(file_path.c_str(), normal.data.data(), normal.cols, normal.rows, 3);
}

hpm_camera ReadCamera(const std::string& cam_path)
{
    hpm_camera c;
    if (hpm_read_camera(cam_path.c_str(), &c) != 0) {
        throw Error("cannot read camera " + cam_path);
    }
    return c;
}

HPM::HPM() : h_(hpm_create(0))
{
    if (!h_) {
        throw Error("hpm_create failed: no HIP device");
    }
    hpm_params_default(&params_);
    memset(&ref_cam_, 0, sizeof(ref_cam_));
}

HPM::~HPM()
{
    hpm_destroy(h_);
}

void HPM::check(int rc, const char* what)
{
    if (rc < 0) {
        throw Error(std::string(what) + ": " + hpm_last_error(h_));
    }
}

void HPM::SetGeomConsistencyParams(bool multi_geometry)
{
    check(hpm_set_geom_consistency(h_, multi_geometry ? 1 : 0), "SetGeomConsistencyParams");
    hpm_get_params(h_, &params_);
}

void HPM::SetPlanarPriorParams()
{
    check(hpm_set_planar_prior(h_), "SetPlanarPriorParams");
    hpm_get_params(h_, &params_);
}

void HPM::InuputInitialization(const std::string& dense_folder, const Problem& problem)
{
    hpm_params p = params_;
    hpm_params cur;
    hpm_get_params(h_, &cur);
    p.geom_consistency = cur.geom_consistency;
    p.multi_geometry = cur.multi_geometry;
    p.planar_prior = cur.planar_prior;
    check(hpm_set_params(h_, &p), "InuputInitialization");
    check(hpm_input_init(h_, dense_folder.c_str(), problem.ref_image_id, problem.src_image_ids.data(),
                         (int)problem.src_image_ids.size()),
          "InuputInitialization");
    check(hpm_get_camera(h_, 0, &ref_cam_), "InuputInitialization");
    hpm_get_params(h_, &params_);
}

void HPM::CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem)
{
    (void)problem;
    check(hpm_device_init(h_, dense_folder.c_str()), "CudaSpaceInitialization");
}

void HPM::RunPatchMatch()
{
    check(hpm_run_patchmatch(h_), "RunPatchMatch");
}

int HPM::GetReferenceImageWidth()
{
    return ref_cam_.width;
}

int HPM::GetReferenceImageHeight()
{
    return ref_cam_.height;
}

MatF HPM::GetReferenceImage()
{
    MatF m(ref_cam_.height, ref_cam_.width);
    check(hpm_copy_reference_image(h_, m.data.data()), "GetReferenceImage");
    return m;
}

float4 HPM::GetPlaneHypothesis(const int index)
{
    float4 v;
    check(hpm_get_plane(h_, index, &v.x), "GetPlaneHypothesis");
    return v;
}

float HPM::GetCost(const int index)
{
    int w = 0, hh = 0;
    check(hpm_dims(h_, &w, &hh), "GetCost");
    if (index < 0 || (int64_t)index >= (int64_t)w * hh) {
        throw Error("GetCost: pixel index " + std::to_string(index) + " outside the reference image");
    }
    return hpm_get_cost(h_, index);
}

float HPM::GetMinDepth()
{
    return params_.depth_min;
}

float HPM::GetMaxDepth()
{
    return params_.depth_max;
}

void HPM::support_points_of(const MatF& costs, std::vector<Point>& out)
{
    std::vector<int32_t> pts((size_t)(costs.cols / 5 + 1) * (costs.rows / 5 + 1) * 2);
    const int n = hpm_support_points(costs.data.data(), costs.cols, costs.rows, pts.data(), (int)(pts.size() / 2));
    out.clear();
    for (int i = 0; i < n; ++i) {
        out.push_back({pts[2 * i], pts[2 * i + 1]});
    }
}

void HPM::GetSupportPoints(std::vector<Point>& support2DPoints)
{
    MatF costs(ref_cam_.height, ref_cam_.width);
    check(hpm_copy_costs(h_, costs.data.data()), "GetSupportPoints");
    support_points_of(costs, support2DPoints);
}

void HPM::GetSupportPointsScale0(std::vector<Point>& support2DPoints, MatF costs)
{
    support_points_of(costs, support2DPoints);
}

void HPM::GetSupportPointsScale1(std::vector<Point>& support2DPoints, MatF costs)
{
    support_points_of(costs, support2DPoints);
}

std::vector<Triangle> HPM::DelaunayTriangulation(const Rect boundRC, const std::vector<Point>& points)
{
    std::vector<Triangle> out;
    if (points.empty()) {
        return out;
    }
    std::vector<int32_t> pts;
    for (const Point& p : points) {
        pts.push_back(p.x);
        pts.push_back(p.y);
    }
    const int n = hpm_delaunay(pts.data(), (int)points.size(), nullptr, 0);
    std::vector<int32_t> tris((size_t)std::max(n, 0) * 3);
    hpm_delaunay(pts.data(), (int)points.size(), tris.data(), n);
    for (int t = 0; t < n; ++t) {
        const Point a = points[(size_t)tris[3 * t]], b = points[(size_t)tris[3 * t + 1]],
                    c = points[(size_t)tris[3 * t + 2]];
        if (boundRC.contains(a) && boundRC.contains(b) && boundRC.contains(c)) {
            out.emplace_back(a, b, c);
        }
    }
    return out;
}

float4 HPM::GetPriorPlaneParams_factor(const Triangle triangle, const MatF& depths, float factor)
{
    const int32_t p[6] = {triangle.pt1.x, triangle.pt1.y, triangle.pt2.x, triangle.pt2.y, triangle.pt3.x,
                          triangle.pt3.y};
    float4 v;
    hpm_plane_from_points(&ref_cam_, p, depths.data.data(), depths.cols, factor, &v.x);
    return v;
}

float HPM::GetDepthFromPlaneParam_factor(const float4 plane_hypothesis, const int x, const int y, float factor)
{
    return hpm_depth_from_plane_factor(&ref_cam_, &plane_hypothesis.x, x, y, factor);
}

int HPM::ComputingMultiScaleSettings(int width, int height)
{
    return hpm_multiscale_count(width, height);
}

void HPM::WritePointCloud(PointCloudPtr& cloud, std::vector<Point> points)
{
    for (const Point& p : points) {
        cloud->push_back(p);
    }
}

/* HPM.cpp:872-890 */
float4 HPM::TransformNormal(float4 n)
{
    const hpm_camera& c = ref_cam_;
    float4 t;
    t.x = c.R[0] * n.x + c.R[3] * n.y + c.R[6] * n.z;
    t.y = c.R[1] * n.x + c.R[4] * n.y + c.R[7] * n.z;
    t.z = c.R[2] * n.x + c.R[5] * n.y + c.R[8] * n.z;
    t.w = n.w;
    return t;
}

float4 HPM::TransformNormal2RefCam(float4 n)
{
    const hpm_camera& c = ref_cam_;
    float4 t;
    t.x = c.R[0] * n.x + c.R[1] * n.y + c.R[2] * n.z;
    t.y = c.R[3] * n.x + c.R[4] * n.y + c.R[5] * n.z;
    t.z = c.R[6] * n.x + c.R[7] * n.y + c.R[8] * n.z;
    t.w = n.w;
    return t;
}

/* HPM.cpp:892-899: the full-resolution K whatever the factor */
float HPM::GetDistance2Origin_factor(const int2 p, const float depth, const float4 normal, float factor)
{
    (void)factor;
    const hpm_camera& c = ref_cam_;
    const float X0 = depth * ((float)p.x - c.K[2] * 1) / (c.K[0] * 1);
    const float X1 = depth * ((float)p.y - c.K[5] * 1) / (c.K[4] * 1);
    const float X2 = depth;
    return -(normal.x * X0 + normal.y * X1 + normal.z * X2);
}

void HPM::CudaPlanarPriorInitialization(const std::vector<float4>& PlaneParams, const MatF& masks)
{
    const size_t np = masks.data.size();
    std::vector<float> planes(np * 4, 0.0f);
    std::vector<uint32_t> m(np, 0u);
    for (size_t i = 0; i < np; ++i) {
        const float mk = masks.data[i];
        if (mk > 0) {
            m[i] = (uint32_t)mk;
            memcpy(&planes[4 * i], &PlaneParams[(size_t)mk - 1], 16);
        }
    }
    check(hpm_upload_prior(h_, planes.data(), m.data()), "CudaPlanarPriorInitialization");
}

void HPM::ReloadPlanarPriorInitialization(const MatF& masks, float4* prior_plane_parameters)
{
    const size_t np = masks.data.size();
    std::vector<uint32_t> m(np, 0u);
    for (size_t i = 0; i < np; ++i) {
        m[i] = masks.data[i] > 0 ? (uint32_t)masks.data[i] : 0u;
    }
    check(hpm_upload_prior(h_, &prior_plane_parameters[0].x, m.data()), "ReloadPlanarPriorInitialization");
}

void HPM::DepthsPredictSupplement(PointCloudPtr cloud, MatF& depth, MatV3& normal_reliable, float4* plane_parameters,
                                  MatF& mask, float factor)
{
    (void)normal_reliable;
    std::vector<int32_t> pts;
    for (const Point& p : cloud->points) {
        pts.push_back(p.x);
        pts.push_back(p.y);
    }
    check(hpm_knn_supplement(&ref_cam_, pts.data(), (int)cloud->points.size(), depth.data.data(), depth.cols,
                             depth.rows, factor, params_.knn_k, GetMinDepth(), GetMaxDepth(), mask.data.data(),
                             &plane_parameters[0].x, 0),
          "DepthsPredictSupplement");
}

void HPM::CudaPlanarPriorInitializationSupplement(const std::vector<float4>& PlaneParams, MatF& masks,
                                                  const float4* plane_hypothesis)
{
    const size_t np = masks.data.size();
    std::vector<float> planes(np * 4, 0.0f);
    std::vector<uint32_t> m(np, 0u);
    for (size_t i = 0; i < np; ++i) {
        const float mk = masks.data[i];
        if (mk > 0) {
            m[i] = (uint32_t)mk;
            memcpy(&planes[4 * i], &PlaneParams[(size_t)mk - 1], 16);
        } else if (mk < -200.0f) {
            /* (unsigned)(-233.0f) is undefined; the prior is taken as active (Q12) */
            m[i] = 1u;
            memcpy(&planes[4 * i], &plane_hypothesis[i], 16);
            masks.data[i] = 1;
        }
    }
    check(hpm_upload_prior(h_, planes.data(), m.data()), "CudaPlanarPriorInitializationSupplement");
}

void HPM::JointBilateralUpsampling(const MatF& scaled_image_float, const MatF& src_depthmap, MatF& upsample_depthmap,
                                   const MatV3& src_normal, MatV3& upsample_normal)
{
    const int w = scaled_image_float.cols, h = scaled_image_float.rows;
    const int sw = src_depthmap.cols, sh = src_depthmap.rows;
    std::vector<float> src((size_t)sw * sh * 4);
    for (size_t i = 0; i < (size_t)sw * sh; ++i) {
        src[4 * i + 0] = src_normal.data[3 * i + 0];
        src[4 * i + 1] = src_normal.data[3 * i + 1];
        src[4 * i + 2] = src_normal.data[3 * i + 2];
        src[4 * i + 3] = src_depthmap.data[i];
    }
    std::vector<float> od((size_t)w * h), on((size_t)w * h * 4);
    const int rc = hpm_jbu(h_, scaled_image_float.data.data(), w, h, src.data(), sw, sh, od.data(), on.data());
    check(rc, "JointBilateralUpsampling");
    if (rc == 1) {
        return; /* integer scale 1 (HPM.cpp:1219-1222) */
    }
    upsample_depthmap = MatF(h, w);
    upsample_normal = MatV3(h, w);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        upsample_depthmap.data[i] = od[i];
        upsample_normal.data[3 * i + 0] = on[4 * i + 0];
        upsample_normal.data[3 * i + 1] = on[4 * i + 1];
        upsample_normal.data[3 * i + 2] = on[4 * i + 2];
    }
}

void HPM::ReleasePriorCudaMemory()
{
    hpm_release_prior(h_);
}

void HPM::ReleaseProblemCudaMemory()
{
    hpm_release(h_);
}

void HPM::ReleaseProblemHostMemory() {}

/* ------------------------------------------------------------------ ACMM */

namespace {

std::string acmm_dir(const std::string& dense_folder, int id)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "/ACMM_NESP/2333_%08d", id);
    return dense_folder + buf;
}

/* images/%08d.jpg (or .pgm) as grey and cams/%08d_cam.txt */
void read_view(const std::string& dense_folder, int id, MatF& img, hpm_camera& cam)
{
    char path[4096];
    int w = 0, h = 0;
    snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder.c_str(), id);
    if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
        snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder.c_str(), id);
        if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
            throw Error(std::string("cannot read image ") + path);
        }
    }
    img = MatF(h, w);
    if (hpm_read_image_grey(path, img.data.data(), &w, &h) != 0) {
        throw Error(std::string("cannot read image ") + path);
    }
    snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder.c_str(), id);
    cam = ReadCamera(path);
    cam.width = w;
    cam.height = h;
}

} // namespace

ACMM::ACMM() : h_(hpm_create(0))
{
    if (!h_) {
        throw Error("hpm_create failed: no HIP device");
    }
    hpm_params_default(&params_);
    params_.acmm = 1;
    memset(&ref_cam_, 0, sizeof(ref_cam_));
}

ACMM::~ACMM()
{
    hpm_destroy(h_);
}

void ACMM::check(int rc, const char* what)
{
    if (rc < 0) {
        throw Error(std::string(what) + ": " + hpm_last_error(h_));
    }
}

/* ACMM.cpp:434-441 */
void ACMM::SetGeomConsistencyParams(bool multi_geometry)
{
    params_.geom_consistency = 1;
    params_.max_iterations = 2;
    if (multi_geometry) {
        params_.multi_geometry = 1;
    }
}

/* ACMM.cpp:443-446 */
void ACMM::SetHierarchyParams()
{
    hierarchy_ = true;
}

/* ACMM.cpp:448-553: images scaled to the level's size of their own
 * problem (problems are indexed by image id), cameras with them; with
 * geometric consistency the depth maps of every image */
void ACMM::InuputInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, const int idx)
{
    if (idx < 0 || (size_t)idx >= problems.size()) {
        throw Error("InuputInitialization: problem index outside the list");
    }
    const Problem& problem = problems[(size_t)idx];
    std::vector<int> ids = {problem.ref_image_id};
    ids.insert(ids.end(), problem.src_image_ids.begin(), problem.src_image_ids.end());
    const size_t n = ids.size();
    std::vector<MatF> imgs(n);
    std::vector<hpm_camera> cams(n);
    std::vector<MatF> depths;
    for (size_t i = 0; i < n; ++i) {
        read_view(dense_folder, ids[i], imgs[i], cams[i]);
        /* ACMM.cpp:487-491 indexes problems by the source's image id */
        if (ids[i] < 0 || (size_t)ids[i] >= problems.size()) {
            throw Error("image id outside the problem list");
        }
        const int cur = i == 0 ? problems[(size_t)idx].cur_image_size : problems[(size_t)ids[i]].cur_image_size;
        const int W = imgs[i].cols, H = imgs[i].rows;
        if (W <= cur && H <= cur) {
            continue;
        }
        const float f = std::min((float)cur / W, (float)cur / H);
        const int nw = (int)std::round(W * f), nh = (int)std::round(H * f);
        MatF o(nh, nw);
        hpm_resize_bilinear(imgs[i].data.data(), W, H, o.data.data(), nw, nh);
        imgs[i] = o;
        const float sx = nw / (float)W, sy = nh / (float)H;
        cams[i].K[0] *= sx;
        cams[i].K[2] *= sx;
        cams[i].K[4] *= sy;
        cams[i].K[5] *= sy;
        cams[i].width = nw;
        cams[i].height = nh;
    }
    if (params_.geom_consistency) {
        depths.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const std::string p = acmm_dir(dense_folder, ids[i]) +
                                  (params_.multi_geometry ? "/depths_geom.dmb" : "/depths.dmb");
            if (readDepthDmb(p, depths[i]) != 0) {
                throw Error("cannot read " + p);
            }
            if (depths[i].cols != cams[i].width || depths[i].rows != cams[i].height) {
                throw Error(p + " is not the size of the view's image");
            }
        }
    }
    std::vector<const float*> ip(n), dp(n);
    for (size_t i = 0; i < n; ++i) {
        ip[i] = imgs[i].data.data();
        dp[i] = depths.empty() ? nullptr : depths[i].data.data();
    }
    check(hpm_set_params(h_, &params_), "InuputInitialization");
    check(hpm_input_init_memory(h_, problem.ref_image_id, (int)n, cams.data(), ip.data(),
                                depths.empty() ? nullptr : dp.data()),
          "InuputInitialization");
    check(hpm_get_camera(h_, 0, &ref_cam_), "InuputInitialization");
}

/* ACMM.cpp:555-719: device buffers; with geometric consistency the view's
 * own depth (depths.dmb or depths_geom.dmb), normals.dmb and costs.dmb as
 * its starting state; with the hierarchy the coarser level's maps */
void ACMM::CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem)
{
    check(hpm_device_init(h_, params_.geom_consistency ? dense_folder.c_str() : nullptr), "CudaSpaceInitialization");
    const std::string dir = acmm_dir(dense_folder, problem.ref_image_id);
    if (hierarchy_) {
        MatF depth;
        MatV3 normal;
        if (readDepthDmb(dir + "/depths_geom.dmb", depth) != 0 || readNormalDmb(dir + "/normals.dmb", normal) != 0 ||
            depth.cols != normal.cols || depth.rows != normal.rows) {
            throw Error("cannot read the coarser maps of " + dir);
        }
        std::vector<float> c4(depth.data.size() * 4);
        for (size_t k = 0; k < depth.data.size(); ++k) {
            c4[4 * k] = normal.data[3 * k];
            c4[4 * k + 1] = normal.data[3 * k + 1];
            c4[4 * k + 2] = normal.data[3 * k + 2];
            c4[4 * k + 3] = depth.data[k];
        }
        check(hpm_set_hierarchy_input(h_, c4.data(), depth.cols, depth.rows), "CudaSpaceInitialization");
    }
}

void ACMM::RunPatchMatch()
{
    check(hpm_run_patchmatch(h_), "RunPatchMatch");
}

int ACMM::GetReferenceImageWidth()
{
    return ref_cam_.width;
}

int ACMM::GetReferenceImageHeight()
{
    return ref_cam_.height;
}

MatF ACMM::GetReferenceImage()
{
    MatF m(ref_cam_.height, ref_cam_.width);
    check(hpm_copy_reference_image(h_, m.data.data()), "GetReferenceImage");
    return m;
}

float4 ACMM::GetPlaneHypothesis(const int index)
{
    float v[4];
    check(hpm_get_plane(h_, index, v), "GetPlaneHypothesis");
    return float4{v[0], v[1], v[2], v[3]};
}

float ACMM::GetCost(const int index)
{
    int w = 0, hh = 0;
    check(hpm_dims(h_, &w, &hh), "GetCost");
    if (index < 0 || (int64_t)index >= (int64_t)w * hh) {
        throw Error("GetCost: pixel index " + std::to_string(index) + " outside the reference image");
    }
    return hpm_get_cost(h_, index);
}

} // namespace hpm
