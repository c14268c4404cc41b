/*
 * hpm_main.cpp — the `HPM <dense_folder>` executable (HPM-MVS/main.cpp:636-677).
 *
 * Default: the device pass schedule of libhpm_mvs (hpm_pipeline_run_folder).
 * --class-api: the reference's ProcessProblem (main.cpp:35-499) written
 * against the drop-in `hpm::HPM` class, file for file (.dmb between passes),
 * to show the class API driving the same path as main.cpp does.
 *
 * Options: --geom-iterations N (2), --knn-k K (6), --no-prior, --no-geom,
 * --learned-prior FILE (device schedule only: the U-Net weight file of the
 * learned prior, which then replaces the planar prior passes where its judge
 * keeps enough pixels), --no-fusion (skip RunFusion, main.cpp:675, which
 * otherwise writes HPM/HPM_model.ply).
 *
 * --acmm: the ACMM+NESP executable instead (ACMM+NESP/windows/main.cpp:
 * 302-378, hpm_acmm_run_folder): coarse-to-fine levels with the ACMM kernel
 * variant, results in ACMM_NESP/, fusion to ACMM_NESP/ACMM_NESP_model.ply;
 * --geom-iterations N sets its geometric passes per level (2), --scale-bound
 * N the size bound of its levels (1000). With --class-api the same schedule
 * runs as the reference's main and ProcessProblem do (main.cpp:74-129,
 * 302-378), against the drop-in `hpm::ACMM` class, file for file.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <vector>

#include "HPM.h"
#include "hpm_io.h"
#include "hpm_fusion.h"
#include "hpm_pipeline.h"
#include "hpm_acmm.h"

using namespace hpm;

namespace {

struct Options {
    int geom_iterations = 2;
    int knn_k = 6;
    bool planar_prior = true;
    bool geom = true;
    bool class_api = false;
};

std::vector<Problem> GenerateSampleList(const std::string& dense_folder)
{
    const std::string path = dense_folder + "/pair.txt";
    int n = 0;
    std::vector<Problem> problems;
    if (hpm_read_pair(path.c_str(), &n, nullptr, nullptr, nullptr, 0) != 0) {
        return problems;
    }
    std::vector<int> refs((size_t)n), ns((size_t)n), srcs((size_t)n * HPM_MAX_SOURCES);
    hpm_read_pair(path.c_str(), &n, refs.data(), ns.data(), srcs.data(), HPM_MAX_SOURCES);
    for (int i = 0; i < n; ++i) {
        Problem p;
        p.ref_image_id = refs[(size_t)i];
        for (int j = 0; j < ns[(size_t)i]; ++j) {
            p.src_image_ids.push_back(srcs[(size_t)i * HPM_MAX_SOURCES + j]);
        }
        problems.push_back(p);
    }
    return problems;
}

MatF resize(const MatF& m, int w, int h)
{
    MatF o(h, w);
    hpm_resize_bilinear_c(m.data.data(), m.cols, m.rows, 1, o.data.data(), w, h);
    return o;
}

MatV3 resize(const MatV3& m, int w, int h)
{
    MatV3 o(h, w);
    hpm_resize_bilinear_c(m.data.data(), m.cols, m.rows, 3, o.data.data(), w, h);
    return o;
}

/* rasterise triangles into mask = idx + 1 and fit their planes
 * (main.cpp:96-116) */
void rasterize(HPM& hpm, const std::vector<Triangle>& triangles, const Rect& rc, const MatF& depths, float factor,
               MatF& mask, std::vector<float4>& planes)
{
    std::vector<int32_t> pts, tris;
    for (const Triangle& t : triangles) {
        if (!(rc.contains(t.pt1) && rc.contains(t.pt2) && rc.contains(t.pt3))) {
            continue;
        }
        const int base = (int)pts.size() / 2;
        for (const Point& p : {t.pt1, t.pt2, t.pt3}) {
            pts.push_back(p.x);
            pts.push_back(p.y);
        }
        tris.push_back(base);
        tris.push_back(base + 1);
        tris.push_back(base + 2);
        planes.push_back(hpm.GetPriorPlaneParams_factor(t, depths, factor));
    }
    hpm_rasterize_triangles(pts.data(), tris.data(), (int)tris.size() / 3, mask.data.data(), mask.cols, mask.rows);
}

/* ProcessProblem (main.cpp:35-499) */
void ProcessProblem(const std::string& dense_folder, const Problem& problem, bool geom_consistency, bool planar_prior,
                    bool multi_geometry, const Options& opt, uint32_t stream = 0)
{
    printf("Processing image %08d...\n", problem.ref_image_id);
    char buf[4096];
    snprintf(buf, sizeof(buf), "%s/HPM/2333_%08d", dense_folder.c_str(), problem.ref_image_id);
    const std::string result_folder(buf);
    mkdir(result_folder.c_str(), 0755);

    HPM hpm;
    hpm.params().knn_k = opt.knn_k;
    if (geom_consistency) {
        hpm.SetGeomConsistencyParams(multi_geometry);
    }
    hpm.InuputInitialization(dense_folder, problem);
    hpm.CudaSpaceInitialization(dense_folder, problem);
    /* the Philox stream hpm_pipeline_run gives the pass (the later prior
     * passes of this object follow it) */
    if (hpm_set_stream_index(hpm.handle(), stream) != 0) {
        throw Error(std::string("hpm_set_stream_index: ") + hpm_last_error(hpm.handle()));
    }
    hpm.RunPatchMatch();

    const int width = hpm.GetReferenceImageWidth();
    const int height = hpm.GetReferenceImageHeight();
    MatF depths(height, width), costs(height, width);
    MatV3 normals(height, width);
    auto readback = [&]() {
        for (int col = 0; col < width; ++col) {
            for (int row = 0; row < height; ++row) {
                const int center = row * width + col;
                const float4 ph = hpm.GetPlaneHypothesis(center);
                depths(row, col) = ph.w;
                float* n = normals(row, col);
                n[0] = ph.x;
                n[1] = ph.y;
                n[2] = ph.z;
                costs(row, col) = hpm.GetCost(center);
            }
        }
    };
    readback();

    if (planar_prior) {
        hpm.SetPlanarPriorParams();
        const int scale = hpm.ComputingMultiScaleSettings(width, height);
        PointCloud cloud;
        PointCloudPtr pointcloud = &cloud;
        for (int i = 2 - scale; i <= 2; i++) {
            if (i == 0 || i == 1) {
                const float f = i == 0 ? 0.25f : 0.5f;
                const int ws = (int)lround(width * (double)f), hs = (int)lround(height * (double)f);
                MatF depths_s = resize(depths, ws, hs), costs_s = resize(costs, ws, hs);
                MatV3 normals_s = resize(normals, ws, hs);
                std::vector<Point> support;
                if (i == 0) {
                    hpm.GetSupportPointsScale0(support, costs_s);
                } else {
                    hpm.GetSupportPointsScale1(support, costs_s);
                }
                const Rect rc{0, 0, ws, hs};
                const auto triangles = hpm.DelaunayTriangulation(rc, support);
                MatF mask(hs, ws);
                std::vector<float4> planes;
                rasterize(hpm, triangles, rc, depths_s, f, mask, planes);
                std::vector<float4> supplement((size_t)ws * hs);
                pointcloud->clear();
                hpm.WritePointCloud(pointcloud, support);
                hpm.DepthsPredictSupplement(pointcloud, depths_s, normals_s, supplement.data(), mask, f);
                MatF pd(hs, ws);
                MatV3 pn(hs, ws);
                for (int x = 0; x < ws; ++x) {
                    for (int y = 0; y < hs; ++y) {
                        float4 pl;
                        bool ok = false;
                        if (mask(y, x) > 0) {
                            pl = planes[(size_t)mask(y, x) - 1];
                            const float d = hpm.GetDepthFromPlaneParam_factor(pl, x, y, f);
                            ok = d <= hpm.GetMaxDepth() * 1.2f && d >= hpm.GetMinDepth() * 0.6f;
                            if (ok) pd(y, x) = d;
                        } else if (mask(y, x) < -200) {
                            pl = supplement[(size_t)y * ws + x];
                            const float d = hpm.GetDepthFromPlaneParam_factor(pl, x, y, f);
                            ok = d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth();
                            if (ok) pd(y, x) = d;
                        }
                        if (ok) {
                            const float4 t = hpm.TransformNormal(pl);
                            float* n = pn(y, x);
                            n[0] = t.x;
                            n[1] = t.y;
                            n[2] = t.z;
                        } else {
                            mask(y, x) = 0;
                        }
                    }
                }
                const MatF image = hpm.GetReferenceImage();
                MatF up_d(height, width);
                MatV3 up_n(height, width);
                hpm.JointBilateralUpsampling(image, pd, up_d, pn, up_n);
                MatF up_mask(height, width);
                std::vector<float4> prior((size_t)width * height);
                for (int x = 0; x < width; x++) {
                    for (int y = 0; y < height; y++) {
                        const float d = up_d(y, x);
                        if (d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth()) {
                            up_mask(y, x) = 1;
                            const float* n = up_n(y, x);
                            float4 t = hpm.TransformNormal2RefCam({n[0], n[1], n[2], d});
                            t.w = hpm.GetDistance2Origin_factor({x, y}, d, t, f);
                            prior[(size_t)y * width + x] = t;
                        }
                    }
                }
                hpm.ReloadPlanarPriorInitialization(up_mask, prior.data());
            } else {
                std::vector<Point> support;
                hpm.GetSupportPoints(support);
                const Rect rc{0, 0, width, height};
                const auto triangles = hpm.DelaunayTriangulation(rc, support);
                MatF mask(height, width);
                std::vector<float4> planes;
                rasterize(hpm, triangles, rc, depths, 1.0f, mask, planes);
                std::vector<float4> supplement((size_t)width * height);
                pointcloud->clear();
                hpm.WritePointCloud(pointcloud, support);
                hpm.DepthsPredictSupplement(pointcloud, depths, normals, supplement.data(), mask, 1.0f);
                MatF prior_depths(height, width);
                for (int x = 0; x < width; ++x) {
                    for (int y = 0; y < height; ++y) {
                        const float mk = mask(y, x);
                        if (mk > 0 || mk < -200) {
                            const float4 pl = mk > 0 ? planes[(size_t)mk - 1] : supplement[(size_t)y * width + x];
                            const float d = hpm.GetDepthFromPlaneParam_factor(pl, x, y, 1.0f);
                            if (d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth()) {
                                prior_depths(y, x) = d;
                            } else {
                                mask(y, x) = 0;
                            }
                        } else {
                            mask(y, x) = 0;
                        }
                    }
                }
                writeDepthDmb(result_folder + "/depths_prior.dmb", prior_depths);
                hpm.CudaPlanarPriorInitializationSupplement(planes, mask, supplement.data());
            }
            hpm.RunPatchMatch();
            hpm.ReleasePriorCudaMemory();
            readback();
        }
    }
    const std::string suffix = geom_consistency ? "/depths_geom.dmb" : "/depths.dmb";
    writeDepthDmb(result_folder + suffix, depths);
    writeNormalDmb(result_folder + "/normals.dmb", normals);
    writeDepthDmb(result_folder + "/costs.dmb", costs);
    hpm.ReleaseProblemCudaMemory();
    hpm.ReleaseProblemHostMemory();
    printf("Processing image %08d done!\n", problem.ref_image_id);
}

/* ComputeMultiScaleSettings (ACMM+NESP/windows/main.cpp:35-71) */
int ComputeMultiScaleSettings(const std::string& dense_folder, std::vector<Problem>& problems, int size_bound)
{
    hpm_params defaults;
    hpm_params_default(&defaults);
    int max_num_downscale = -1;
    char path[4096];
    for (Problem& p : problems) {
        int w = 0, h = 0;
        snprintf(path, sizeof(path), "%s/images/%08d.jpg", dense_folder.c_str(), p.ref_image_id);
        if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
            snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder.c_str(), p.ref_image_id);
            if (hpm_read_image_grey(path, nullptr, &w, &h) != 0) {
                throw Error(std::string("cannot read image ") + path);
            }
        }
        int max_size = std::min(std::max(w, h), (int)defaults.max_image_size);
        p.max_image_size = max_size;
        int k = 0;
        while (max_size > size_bound) {
            max_size /= 2;
            ++k;
        }
        max_num_downscale = std::max(max_num_downscale, k);
        p.num_downscale = k;
    }
    return max_num_downscale;
}

/* ProcessProblem of ACMM+NESP (ACMM+NESP/windows/main.cpp:74-129) */
void AcmmProcessProblem(const std::string& dense_folder, const std::vector<Problem>& problems, int idx,
                        uint32_t stream, bool geom_consistency, bool hierarchy, bool multi_geometry = false)
{
    const Problem& problem = problems[(size_t)idx];
    printf("Processing image %08d...\n", problem.ref_image_id);
    char buf[4096];
    snprintf(buf, sizeof(buf), "%s/ACMM_NESP/2333_%08d", dense_folder.c_str(), problem.ref_image_id);
    const std::string result_folder(buf);
    mkdir(result_folder.c_str(), 0755);

    ACMM acmm;
    if (geom_consistency) {
        acmm.SetGeomConsistencyParams(multi_geometry);
    }
    if (hierarchy) {
        acmm.SetHierarchyParams();
    }
    acmm.InuputInitialization(dense_folder, problems, idx);
    acmm.CudaSpaceInitialization(dense_folder, problem);
    /* the Philox stream hpm_acmm_run gives this pass, so that both runs of
     * the schedule draw the same numbers */
    if (hpm_set_stream_index(acmm.handle(), stream) != 0) {
        throw Error(std::string("hpm_set_stream_index: ") + hpm_last_error(acmm.handle()));
    }
    acmm.RunPatchMatch();

    const int width = acmm.GetReferenceImageWidth();
    const int height = acmm.GetReferenceImageHeight();
    MatF depths(height, width), costs(height, width);
    MatV3 normals(height, width);
    for (int col = 0; col < width; ++col) {
        for (int row = 0; row < height; ++row) {
            const int center = row * width + col;
            const float4 plane = acmm.GetPlaneHypothesis(center);
            depths.data[(size_t)center] = plane.w;
            normals.data[3 * (size_t)center] = plane.x;
            normals.data[3 * (size_t)center + 1] = plane.y;
            normals.data[3 * (size_t)center + 2] = plane.z;
            costs.data[(size_t)center] = acmm.GetCost(center);
        }
    }
    const std::string suffix = geom_consistency ? "/depths_geom.dmb" : "/depths.dmb";
    if (writeDepthDmb(result_folder + suffix, depths) != 0 ||
        writeNormalDmb(result_folder + "/normals.dmb", normals) != 0 ||
        writeDepthDmb(result_folder + "/costs.dmb", costs) != 0) {
        throw Error("cannot write the maps of " + result_folder);
    }
    printf("Processing image %08d done!\n", problem.ref_image_id);
}

/* main of ACMM+NESP (ACMM+NESP/windows/main.cpp:302-378) against the class;
 * the JointBilateralUpsampling file step is folded into the hierarchy pass,
 * which upsamples the coarser level's maps itself */
int RunAcmmClassApi(const std::string& dense_folder, int geom_iterations, int size_bound, bool fusion)
{
    std::vector<Problem> problems = GenerateSampleList(dense_folder);
    if (problems.empty()) {
        throw Error("cannot read " + dense_folder + "/pair.txt");
    }
    mkdir((dense_folder + "/ACMM_NESP").c_str(), 0755);
    const size_t num_images = problems.size();
    printf("There are %zu problems needed to be processed!\n", num_images);
    int max_num_downscale = ComputeMultiScaleSettings(dense_folder, problems, size_bound);
    bool first = true;
    for (uint32_t step = 0; max_num_downscale >= 0; ++step) {
        printf("Scale: %d\n", max_num_downscale);
        for (Problem& p : problems) {
            if (p.num_downscale >= 0) {
                p.cur_image_size = (int)(p.max_image_size / pow(2, p.num_downscale));
                p.num_downscale--;
            }
        }
        for (size_t i = 0; i < num_images; ++i) {
            AcmmProcessProblem(dense_folder, problems, (int)i, 16u * step, false, !first);
        }
        for (int g = 0; g < geom_iterations; ++g) {
            for (size_t i = 0; i < num_images; ++i) {
                AcmmProcessProblem(dense_folder, problems, (int)i, 16u * step + 8u + (uint32_t)g, true, false, g > 0);
            }
        }
        first = false;
        max_num_downscale--;
    }
    if (!fusion) {
        return 0;
    }
    int64_t n = 0;
    char err[512] = {0};
    if (hpm_fusion_run_folder_acmm(dense_folder.c_str(), geom_iterations > 0 ? 1 : 0, &n, err, sizeof(err)) != 0) {
        fprintf(stderr, "hpm_mvs: %s\n", err);
        return 1;
    }
    printf("fusion: %lld points -> %s/ACMM_NESP/ACMM_NESP_model.ply\n", (long long)n, dense_folder.c_str());
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        printf("USAGE: hpm_mvs dense_folder [--class-api] [--geom-iterations N] [--knn-k K] [--no-prior] "
               "[--no-geom] [--learned-prior FILE] [--no-fusion] [--acmm] [--scale-bound N]\n");
        return -1;
    }
    const std::string dense_folder = argv[1];
    Options opt;
    const char* learned = nullptr;
    bool fusion = true;
    bool acmm = false;
    int scale_bound = 1000;
    /* RunFusion (main.cpp:675) on the maps just written */
    auto run_fusion = [&]() -> int {
        if (!fusion) {
            return 0;
        }
        int64_t n = 0;
        char ferr[512] = {0};
        if (hpm_fusion_run_folder(dense_folder.c_str(), opt.geom ? 1 : 0, &n, ferr, sizeof(ferr)) != 0) {
            fprintf(stderr, "hpm_mvs: %s\n", ferr);
            return 1;
        }
        printf("fusion: %lld points -> %s/HPM/HPM_model.ply\n", (long long)n, dense_folder.c_str());
        return 0;
    };
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--class-api")) {
            opt.class_api = true;
        } else if (!strcmp(argv[i], "--geom-iterations") && i + 1 < argc) {
            opt.geom_iterations = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--knn-k") && i + 1 < argc) {
            opt.knn_k = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--no-prior")) {
            opt.planar_prior = false;
        } else if (!strcmp(argv[i], "--no-geom")) {
            opt.geom = false;
        } else if (!strcmp(argv[i], "--learned-prior") && i + 1 < argc) {
            learned = argv[++i];
        } else if (!strcmp(argv[i], "--no-fusion")) {
            fusion = false;
        } else if (!strcmp(argv[i], "--acmm")) {
            acmm = true;
        } else if (!strcmp(argv[i], "--scale-bound") && i + 1 < argc) {
            scale_bound = atoi(argv[++i]);
        } else {
            fprintf(stderr, "unknown option %s\n", argv[i]);
            return -1;
        }
    }
    if (acmm) {
        if (learned || scale_bound < 1) {
            fprintf(stderr, "hpm_mvs: --acmm takes neither --learned-prior nor a scale bound below 1\n");
            return -1;
        }
        if (opt.class_api) {
            try {
                return RunAcmmClassApi(dense_folder, opt.geom ? opt.geom_iterations : 0, scale_bound, fusion);
            } catch (const std::exception& e) {
                fprintf(stderr, "hpm_mvs: %s\n", e.what());
                return 1;
            }
        }
        hpm_acmm_options ao;
        hpm_acmm_default(&ao);
        ao.geom_iterations = opt.geom ? opt.geom_iterations : 0;
        ao.scale_bound = scale_bound;
        hpm_pipeline_stats st;
        int64_t n = 0;
        char err[512] = {0};
        if (hpm_acmm_run_folder(dense_folder.c_str(), &ao, fusion ? 1 : 0, &n, &st, err, sizeof(err)) != 0) {
            fprintf(stderr, "hpm_mvs: %s\n", err);
            return 1;
        }
        printf("ACMM+NESP: %d RunPatchMatch calls, %.1f ms device\n", st.runs, st.ms_device);
        if (fusion) {
            printf("fusion: %lld points -> %s/ACMM_NESP/ACMM_NESP_model.ply\n", (long long)n, dense_folder.c_str());
        }
        return 0;
    }
    mkdir((dense_folder + "/HPM").c_str(), 0755);
    if (!opt.class_api) {
        hpm_pipeline_options po;
        hpm_pipeline_default(&po);
        po.params.knn_k = opt.knn_k;
        po.params.geom_iterations = opt.geom_iterations;
        po.planar_prior = opt.planar_prior ? 1 : 0;
        po.geom_consistency = opt.geom ? 1 : 0;
        po.learned_prior = learned;
        hpm_pipeline_stats st;
        char err[512] = {0};
        if (hpm_pipeline_run_folder(dense_folder.c_str(), &po, &st, err, sizeof(err)) != 0) {
            fprintf(stderr, "hpm_mvs: %s\n", err);
            return 1;
        }
        printf("%d RunPatchMatch calls, %.1f ms wall, %.1f ms device, %.1f ms host prior\n", st.runs, st.ms_wall,
               st.ms_device, st.ms_prior_host);
        if (learned) {
            printf("learned prior: %d views kept it, %d fell back to the planar prior\n", st.learned_views,
                   st.learned_fallbacks);
        }
        return run_fusion();
    }
    if (learned) {
        fprintf(stderr, "hpm_mvs: --learned-prior runs with the device schedule, not with --class-api\n");
        return -1;
    }
    try {
        const std::vector<Problem> problems = GenerateSampleList(dense_folder);
        printf("There are %zu problems needed to be processed!\n", problems.size());
        for (const Problem& p : problems) {
            ProcessProblem(dense_folder, p, false, opt.planar_prior, false, opt);
        }
        if (opt.geom) {
            for (int g = 0; g < opt.geom_iterations; ++g) {
                for (const Problem& p : problems) {
                    ProcessProblem(dense_folder, p, true, false, g > 0, opt, 8u + (uint32_t)g);
                }
            }
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "hpm_mvs: %s\n", e.what());
        return 1;
    }
    return run_fusion();
}
