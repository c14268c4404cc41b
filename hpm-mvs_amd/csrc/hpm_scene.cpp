/*
 * hpm_scene.cpp — seeded synthetic scenes with analytic ground truth
 * (include/hpm_scene.h). Ray casting against plane patches and boxes, value
 * noise texture on the surfaces, per-image Gaussian noise.
 */
#include "hpm_scene.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hpm_io.h"
#include "hpm_philox.h"

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a)
{
    const double n = sqrt(dot(a, a));
    return {a.x / n, a.y / n, a.z / n};
}

/* sequential uniform stream for scene parameters */
struct Rng {
    hpm_rng r;
    explicit Rng(uint32_t seed, uint32_t stream) { hpm_rng_init(&r, seed, 0x5ce7e5eeu, 0u, stream, 0x0b1ec75u); }
    double u() { return (double)(hpm_rng_next_u32(&r) >> 8) * (1.0 / 16777216.0); }
    double range(double a, double b) { return a + (b - a) * u(); }
};

enum Kind { BACKGROUND = 0, PLANE = 1, BOX = 2 };

struct Object {
    Kind kind;
    V3 p0;          /* plane anchor / box centre */
    V3 n, u, v;     /* plane normal and in-plane axes */
    double ha, hb;  /* plane half extents */
    double cphi, sphi; /* box rotation about y */
    V3 half;        /* box half sizes */
    double pitch;   /* texture lattice pitch in world units */
    uint32_t id;
    float amp;
};

struct Hit {
    double t;
    int obj;
    int face;
    double s, q; /* surface coordinates */
    V3 n;
};

/* bilinear value noise on a lattice of hashed values in [-1, 1] */
double value_noise(uint32_t seed, uint32_t obj, int face, double s, double q)
{
    const double fs = floor(s), fq = floor(q);
    const double a = s - fs, b = q - fq;
    const int64_t i = (int64_t)fs, j = (int64_t)fq;
    double L[2][2];
    for (int dj = 0; dj < 2; ++dj) {
        for (int di = 0; di < 2; ++di) {
            const uint32_t ctr[4] = {(uint32_t)(i + di), (uint32_t)(j + dj), obj * 8u + (uint32_t)face, seed};
            const uint32_t key[2] = {0x7e47e4u, 0x6e015eu};
            uint32_t out[4];
            hpm_philox4x32_10(ctr, key, out);
            L[dj][di] = (double)(out[0] >> 8) * (2.0 / 16777216.0) - 1.0;
        }
    }
    const double top = L[0][0] + a * (L[0][1] - L[0][0]);
    const double bot = L[1][0] + a * (L[1][1] - L[1][0]);
    return top + b * (bot - top);
}

struct Scene {
    hpm_scene_config cfg;
    std::vector<Object> objs;
    std::vector<hpm_camera> cams;
    std::vector<V3> centres;
    std::vector<double> Rw[1]; /* unused */
};

void make_cameras(Scene* sc)
{
    const hpm_scene_config& c = sc->cfg;
    const int n = c.num_views;
    for (int i = 0; i < n; ++i) {
        const double theta =
            n > 1 ? ((double)i - 0.5 * (n - 1)) * (double)c.arc_length / ((double)c.look_depth * (n - 1)) : 0.0;
        const V3 C = {c.look_depth * sin(theta), 0.0, c.look_depth - c.look_depth * cos(theta)};
        /* camera axes in world: x = (cos, 0, sin), y = (0, 1, 0), z = (-sin, 0, cos) */
        const V3 xa = {cos(theta), 0.0, sin(theta)}, ya = {0.0, 1.0, 0.0}, za = {-sin(theta), 0.0, cos(theta)};
        hpm_camera cam;
        memset(&cam, 0, sizeof(cam));
        const V3 rows[3] = {xa, ya, za};
        for (int r = 0; r < 3; ++r) {
            cam.R[3 * r + 0] = (float)rows[r].x;
            cam.R[3 * r + 1] = (float)rows[r].y;
            cam.R[3 * r + 2] = (float)rows[r].z;
            cam.t[r] = (float)(-dot(rows[r], C));
        }
        cam.K[0] = c.fx;
        cam.K[2] = c.cx;
        cam.K[4] = c.fy;
        cam.K[5] = c.cy;
        cam.K[8] = 1.0f;
        cam.width = c.width;
        cam.height = c.height;
        cam.depth_min = c.depth_min;
        cam.depth_max = c.depth_max;
        sc->cams.push_back(cam);
        sc->centres.push_back(C);
    }
}

void make_objects(Scene* sc)
{
    const hpm_scene_config& c = sc->cfg;
    Rng rng(c.seed, 1u);
    Object bg;
    memset(&bg, 0, sizeof(bg));
    bg.kind = BACKGROUND;
    bg.p0 = {0.0, 0.0, c.depth_max};
    bg.n = {0.0, 0.0, -1.0};
    bg.u = {1.0, 0.0, 0.0};
    bg.v = {0.0, 1.0, 0.0};
    bg.pitch = c.pitch_px * c.depth_max / c.fx;
    bg.id = 0;
    sc->objs.push_back(bg);
    auto anchor = [&](double zlo, double zhi) {
        const double x = rng.range(0.0, c.width), y = rng.range(0.0, c.height);
        const double z = rng.range(zlo, zhi);
        return V3{z * (x - c.cx) / c.fx, z * (y - c.cy) / c.fy, z};
    };
    for (int i = 0; i < c.num_planes; ++i) {
        Object o;
        memset(&o, 0, sizeof(o));
        o.kind = PLANE;
        o.p0 = anchor(c.depth_min + 0.25 * (c.depth_max - c.depth_min), c.depth_max - 0.1 * (c.depth_max - c.depth_min));
        /* normal uniform on the 60-degree cap around -z */
        const double cth = rng.range(0.5, 1.0);
        const double sth = sqrt(1.0 - cth * cth);
        const double phi = rng.range(0.0, 2.0 * M_PI);
        o.n = {sth * cos(phi), sth * sin(phi), -cth};
        V3 ref = fabs(o.n.y) < 0.9 ? V3{0.0, 1.0, 0.0} : V3{1.0, 0.0, 0.0};
        o.u = normalize(cross(o.n, ref));
        o.v = cross(o.n, o.u);
        o.ha = rng.range(0.15, 0.45) * o.p0.z;
        o.hb = rng.range(0.15, 0.45) * o.p0.z;
        o.pitch = c.pitch_px * o.p0.z / c.fx;
        o.id = (uint32_t)sc->objs.size();
        sc->objs.push_back(o);
    }
    for (int i = 0; i < c.num_boxes; ++i) {
        Object o;
        memset(&o, 0, sizeof(o));
        o.kind = BOX;
        o.p0 = anchor(c.depth_min + 0.15 * (c.depth_max - c.depth_min), c.depth_max - 0.2 * (c.depth_max - c.depth_min));
        const double phi = rng.range(0.0, 0.5 * M_PI);
        o.cphi = cos(phi);
        o.sphi = sin(phi);
        o.half = {rng.range(0.03, 0.12) * o.p0.z, rng.range(0.03, 0.12) * o.p0.z, rng.range(0.03, 0.12) * o.p0.z};
        o.pitch = c.pitch_px * o.p0.z / c.fx;
        o.id = (uint32_t)sc->objs.size();
        sc->objs.push_back(o);
    }
}

/* box-local from world: rotate by -phi about y */
inline V3 to_box(const Object& o, V3 w)
{
    return {o.cphi * w.x - o.sphi * w.z, w.y, o.sphi * w.x + o.cphi * w.z};
}
inline V3 from_box(const Object& o, V3 l)
{
    return {o.cphi * l.x + o.sphi * l.z, l.y, -o.sphi * l.x + o.cphi * l.z};
}

bool intersect(const Scene& sc, const Object& o, V3 C, V3 d, Hit* h)
{
    const hpm_scene_config& c = sc.cfg;
    if (o.kind == BACKGROUND || o.kind == PLANE) {
        const double den = dot(d, o.n);
        if (fabs(den) < 1e-12) {
            return false;
        }
        const double t = dot(o.p0 - C, o.n) / den;
        if (!(t > 1e-6)) {
            return false;
        }
        const V3 X = C + t * d;
        const V3 r = X - o.p0;
        const double s = dot(r, o.u), q = dot(r, o.v);
        if (o.kind == PLANE) {
            if (fabs(s) > o.ha || fabs(q) > o.hb || X.z < c.depth_min || X.z > c.depth_max) {
                return false;
            }
        }
        h->t = t;
        h->face = 0;
        h->s = s;
        h->q = q;
        h->n = o.n;
        return true;
    }
    /* box: slab test in box-local coordinates */
    const V3 ol = to_box(o, C - o.p0);
    const V3 dl = to_box(o, d);
    const double oo[3] = {ol.x, ol.y, ol.z}, dd[3] = {dl.x, dl.y, dl.z}, hh[3] = {o.half.x, o.half.y, o.half.z};
    double tn = -1e300, tf = 1e300;
    int axis = -1;
    for (int a = 0; a < 3; ++a) {
        if (fabs(dd[a]) < 1e-15) {
            if (oo[a] < -hh[a] || oo[a] > hh[a]) {
                return false;
            }
            continue;
        }
        double t1 = (-hh[a] - oo[a]) / dd[a], t2 = (hh[a] - oo[a]) / dd[a];
        if (t1 > t2) {
            std::swap(t1, t2);
        }
        if (t1 > tn) {
            tn = t1;
            axis = a;
        }
        tf = std::min(tf, t2);
    }
    if (axis < 0 || tn > tf || !(tn > 1e-6)) {
        return false;
    }
    const V3 X = C + tn * d;
    if (X.z < c.depth_min || X.z > c.depth_max) {
        return false;
    }
    const double pl[3] = {oo[0] + tn * dd[0], oo[1] + tn * dd[1], oo[2] + tn * dd[2]};
    const double sgn = pl[axis] > 0 ? 1.0 : -1.0;
    V3 nl = {0.0, 0.0, 0.0};
    if (axis == 0) nl.x = sgn;
    if (axis == 1) nl.y = sgn;
    if (axis == 2) nl.z = sgn;
    h->t = tn;
    h->face = axis * 2 + (sgn > 0 ? 1 : 0);
    h->s = pl[(axis + 1) % 3];
    h->q = pl[(axis + 2) % 3];
    h->n = from_box(o, nl);
    return true;
}

bool trace(const Scene& sc, V3 C, V3 d, Hit* best)
{
    best->t = 1e300;
    best->obj = -1;
    Hit h;
    for (size_t i = 0; i < sc.objs.size(); ++i) {
        if (intersect(sc, sc.objs[i], C, d, &h) && h.t < best->t) {
            *best = h;
            best->obj = (int)i;
        }
    }
    return best->obj >= 0;
}

/* ray through pixel (x, y) of view v: direction with camera z = 1 */
inline V3 ray_dir(const hpm_camera& cam, double x, double y)
{
    const double lx = (x - cam.K[2]) / cam.K[0], ly = (y - cam.K[5]) / cam.K[4];
    return {cam.R[0] * lx + cam.R[3] * ly + cam.R[6], cam.R[1] * lx + cam.R[4] * ly + cam.R[7],
            cam.R[2] * lx + cam.R[5] * ly + cam.R[8]};
}

double gaussian(uint32_t seed, uint32_t view, uint32_t pixel)
{
    const uint32_t ctr[4] = {pixel, view, 0x6e6f6973u, seed};
    const uint32_t key[2] = {0x9a055u, 0x1a2b3cu};
    uint32_t out[4];
    hpm_philox4x32_10(ctr, key, out);
    const double u1 = ((double)(out[0] >> 8) + 1.0) * (1.0 / 16777216.0);
    const double u2 = (double)(out[1] >> 8) * (1.0 / 16777216.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

int check_cfg(const hpm_scene_config* c)
{
    if (!c || c->width <= 0 || c->height <= 0 || c->num_views < 1 || c->num_views > 512 || c->fx <= 0 ||
        c->fy <= 0 || c->num_planes < 0 || c->num_boxes < 0 || !(c->depth_max > c->depth_min) ||
        c->depth_min <= 0 || c->pitch_px <= 0 || c->look_depth <= 0) {
        return -1;
    }
    return 0;
}

} // namespace

extern "C" void hpm_scene_default(hpm_scene_config* cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    cfg->seed = 1;
    cfg->width = 640;
    cfg->height = 480;
    cfg->fx = 500.0f;
    cfg->fy = 500.0f;
    cfg->cx = 320.0f;
    cfg->cy = 240.0f;
    cfg->num_views = 5;
    cfg->arc_length = 0.3f;
    cfg->look_depth = 5.0f;
    cfg->num_planes = 4;
    cfg->num_boxes = 0;
    cfg->low_texture_fraction = 0.0f;
    cfg->amp_high = 60.0f;
    cfg->amp_low = 3.0f;
    cfg->noise_sigma = 2.0f;
    cfg->pitch_px = 8.0f;
    cfg->depth_min = 2.0f;
    cfg->depth_max = 8.0f;
}

extern "C" int hpm_scene_render(const hpm_scene_config* cfg, hpm_camera* cams, uint8_t* images, float* gt_depth,
                                float* gt_normal, uint8_t* low_texture)
{
    if (check_cfg(cfg) != 0 || !cams || !images) {
        return -1;
    }
    Scene sc;
    sc.cfg = *cfg;
    make_cameras(&sc);
    make_objects(&sc);
    const int W = cfg->width, H = cfg->height, n = cfg->num_views;
    const size_t np = (size_t)W * H;
    const int mid = (n - 1) / 2;

    /* low-texture assignment by middle-view coverage */
    std::vector<int> ids(np, -1);
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            Hit h;
            if (trace(sc, sc.centres[mid], ray_dir(sc.cams[mid], x, y), &h)) {
                ids[(size_t)y * W + x] = h.obj;
            }
        }
    }
    std::vector<size_t> cover(sc.objs.size(), 0);
    for (int id : ids) {
        if (id >= 0) {
            cover[(size_t)id]++;
        }
    }
    std::vector<int> order(sc.objs.size());
    for (size_t i = 0; i < order.size(); ++i) {
        order[i] = (int)i;
    }
    Rng rng(cfg->seed, 2u);
    for (size_t i = order.size(); i > 1; --i) { /* Fisher-Yates */
        const size_t j = (size_t)(rng.u() * (double)i) % i;
        std::swap(order[i - 1], order[j]);
    }
    for (auto& o : sc.objs) {
        o.amp = cfg->amp_high;
    }
    size_t low = 0;
    const size_t target = (size_t)((double)cfg->low_texture_fraction * (double)np);
    for (int idx : order) {
        if (low >= target) {
            break;
        }
        if (cover[(size_t)idx] == 0) {
            continue;
        }
        sc.objs[(size_t)idx].amp = cfg->amp_low;
        low += cover[(size_t)idx];
    }
    if (low_texture) {
        for (size_t i = 0; i < np; ++i) {
            low_texture[i] = (ids[i] >= 0 && sc.objs[(size_t)ids[i]].amp == cfg->amp_low && cfg->amp_low != cfg->amp_high)
                                 ? 1
                                 : 0;
        }
    }

    for (int v = 0; v < n; ++v) {
        cams[v] = sc.cams[v];
        uint8_t* img = images + (size_t)v * np;
        float* gd = gt_depth ? gt_depth + (size_t)v * np : nullptr;
        float* gn = gt_normal ? gt_normal + (size_t)v * np * 3 : nullptr;
#pragma omp parallel for schedule(dynamic, 8)
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                const size_t p = (size_t)y * W + x;
                const V3 d = ray_dir(sc.cams[v], x, y);
                Hit h;
                double val = 128.0;
                float depth = 0.0f;
                V3 nrm = {0.0, 0.0, 0.0};
                if (trace(sc, sc.centres[v], d, &h)) {
                    const Object& o = sc.objs[(size_t)h.obj];
                    val = 128.0 + o.amp * value_noise(cfg->seed, o.id, h.face, h.s / o.pitch, h.q / o.pitch);
                    depth = (float)h.t;
                    nrm = h.n;
                    if (dot(nrm, d) > 0) {
                        nrm = -1.0 * nrm;
                    }
                }
                val += cfg->noise_sigma * gaussian(cfg->seed, (uint32_t)v, (uint32_t)p);
                const double r = floor(val + 0.5);
                img[p] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
                if (gd) {
                    gd[p] = depth;
                }
                if (gn) {
                    gn[3 * p + 0] = (float)nrm.x;
                    gn[3 * p + 1] = (float)nrm.y;
                    gn[3 * p + 2] = (float)nrm.z;
                }
            }
        }
    }
    return 0;
}

extern "C" int hpm_scene_write_folder(const hpm_scene_config* cfg, const char* dense_folder)
{
    if (check_cfg(cfg) != 0 || !dense_folder) {
        return -1;
    }
    const int W = cfg->width, H = cfg->height, n = cfg->num_views;
    const size_t np = (size_t)W * H;
    std::vector<hpm_camera> cams((size_t)n);
    std::vector<uint8_t> imgs(np * n);
    std::vector<float> gd(np * n), gn(np * n * 3);
    if (hpm_scene_render(cfg, cams.data(), imgs.data(), gd.data(), gn.data(), nullptr) != 0) {
        return -1;
    }
    const std::string base(dense_folder);
    for (const char* sub : {"", "/images", "/cams", "/gt"}) {
        mkdir((base + sub).c_str(), 0755);
    }
    for (int v = 0; v < n; ++v) {
        char path[4096];
        snprintf(path, sizeof(path), "%s/images/%08d.pgm", dense_folder, v);
        if (hpm_write_pgm_u8(path, &imgs[np * v], W, H) != 0) return -2;
        snprintf(path, sizeof(path), "%s/cams/%08d_cam.txt", dense_folder, v);
        if (hpm_write_camera(path, &cams[(size_t)v]) != 0) return -2;
        snprintf(path, sizeof(path), "%s/gt/%08d_depth.dmb", dense_folder, v);
        if (hpm_write_dmb(path, &gd[np * v], W, H, 1) != 0) return -2;
        snprintf(path, sizeof(path), "%s/gt/%08d_normal.dmb", dense_folder, v);
        if (hpm_write_dmb(path, &gn[np * 3 * v], W, H, 3) != 0) return -2;
    }
    FILE* f = fopen((base + "/pair.txt").c_str(), "w");
    if (!f) {
        return -2;
    }
    fprintf(f, "%d\n", n);
    for (int v = 0; v < n; ++v) {
        std::vector<int> others;
        for (int u = 0; u < n; ++u) {
            if (u != v) {
                others.push_back(u);
            }
        }
        std::stable_sort(others.begin(), others.end(), [v](int a, int b) { return abs(a - v) < abs(b - v); });
        fprintf(f, "%d\n%d", v, (int)others.size());
        for (int u : others) {
            fprintf(f, " %d %.6f", u, 100.0 / (1.0 + abs(u - v)));
        }
        fprintf(f, "\n");
    }
    return fclose(f) == 0 ? 0 : -2;
}
