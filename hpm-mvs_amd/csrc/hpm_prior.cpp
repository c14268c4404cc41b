/*
 * hpm_prior.cpp — host construction of the planar prior (include/hpm_prior.h).
 */
#include "hpm_prior.h"

#include "hpm_dt.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

/* The host grid of the star walk (hpm_dt.h): cells of cs pixels over the
 * points' bounding box, each listing its points; the convex hull by
 * Andrew's monotone chain. */
struct HostGrid {
    const std::vector<hpm_dt::P2>* P;
    int x0, y0, gw, gh, cs;
    std::vector<int> start, items;
    std::vector<hpm_dt::P2> H;
    int nhull;

    hpm_dt::P2 pt(int i) const { return (*P)[(size_t)i]; }
    int cx(hpm_dt::P2 p) const { return (p.x - x0) / cs; }
    int cy(hpm_dt::P2 p) const { return (p.y - y0) / cs; }
    int cxd(double x) const { return (int)fmax(-1.0, fmin((double)gw, floor((x - x0) / cs))); }
    int cyd(double y) const { return (int)fmax(-1.0, fmin((double)gh, floor((y - y0) / cs))); }
    double rowy(int y) const { return (double)y0 + (double)y * cs; }
    double colx(int x) const { return (double)x0 + (double)x * cs; }
    hpm_dt::P2 hull(int k) const { return H[(size_t)k]; }
    template <class F>
    void visit(int x, int y, F&& f) const
    {
        const size_t c = (size_t)y * gw + x;
        for (int k = start[c]; k < start[c + 1]; ++k) {
            f(items[(size_t)k]);
        }
    }
    template <class F>
    void visit_row(int y, int xa, int xb, F&& f) const
    {
        for (int x = xa; x <= xb; ++x) {
            visit(x, y, f);
        }
    }
    template <class F>
    void visit_col(int x, int ya, int yb, F&& f) const
    {
        for (int y = ya; y <= yb; ++y) {
            visit(x, y, f);
        }
    }

    /* keep: the points to triangulate (distinct coordinates) */
    HostGrid(const std::vector<hpm_dt::P2>& pts, const std::vector<int>& keep)
    {
        P = &pts;
        int x1 = pts[(size_t)keep[0]].x, y1 = pts[(size_t)keep[0]].y;
        x0 = x1;
        y0 = y1;
        for (int i : keep) {
            x0 = std::min(x0, pts[(size_t)i].x);
            y0 = std::min(y0, pts[(size_t)i].y);
            x1 = std::max(x1, pts[(size_t)i].x);
            y1 = std::max(y1, pts[(size_t)i].y);
        }
        /* about one point per cell */
        const double area = ((double)x1 - x0 + 1) * ((double)y1 - y0 + 1);
        cs = std::max(1, (int)ceil(sqrt(area / (double)keep.size())));
        gw = (x1 - x0) / cs + 1;
        gh = (y1 - y0) / cs + 1;
        start.assign((size_t)gw * gh + 1, 0);
        for (int i : keep) {
            start[(size_t)cy(pts[(size_t)i]) * gw + cx(pts[(size_t)i]) + 1] += 1;
        }
        for (size_t c = 0; c < (size_t)gw * gh; ++c) {
            start[c + 1] += start[c];
        }
        items.resize(keep.size());
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i : keep) {
            items[(size_t)fill[(size_t)cy(pts[(size_t)i]) * gw + cx(pts[(size_t)i])]++] = i;
        }
        /* monotone chain over (x, y) */
        std::vector<hpm_dt::P2> s;
        s.reserve(keep.size());
        for (int i : keep) {
            s.push_back(pts[(size_t)i]);
        }
        std::sort(s.begin(), s.end(),
                  [](hpm_dt::P2 a, hpm_dt::P2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
        std::vector<hpm_dt::P2> h(2 * s.size() + 1);
        size_t k = 0;
        for (size_t i = 0; i < s.size(); ++i) {
            while (k >= 2 && hpm_dt::orient(h[k - 2], h[k - 1], s[i]) <= 0) {
                --k;
            }
            h[k++] = s[i];
        }
        for (size_t i = s.size() - 1, t = k + 1; i > 0; --i) {
            while (k >= t && hpm_dt::orient(h[k - 2], h[k - 1], s[i - 1]) <= 0) {
                --k;
            }
            h[k++] = s[i - 1];
        }
        H.assign(h.begin(), h.begin() + (long)(k > 1 ? k - 1 : k));
        nhull = (int)H.size();
    }
};

/* Static 2-D tree over integer points for exact k-nearest-neighbour
 * queries ordered by (squared distance, index), the order the KNN
 * supplement defines (its reference uses a FLANN kd-tree, HPM.cpp:935-947).
 * Subtrees are pruned only when their box lies strictly farther than the
 * current k-th candidate, so equal-distance points of lower index are never
 * missed. */
class KdTree {
public:
    KdTree(const int32_t* pts, int n) : P(pts), idx((size_t)n)
    {
        for (int i = 0; i < n; ++i) {
            idx[i] = i;
        }
        nodes.reserve(2 * (size_t)n / kLeaf + 16);
        if (n > 0) {
            build(0, n);
        }
    }

    /* k nearest of (x, y): squared distances bd and indices bi, ascending */
    int knn(int x, int y, int k, int64_t* bd, int* bi) const
    {
        int found = 0;
        if (nodes.empty()) {
            return 0;
        }
        int stack[64];
        int top = 0;
        stack[top++] = 0;
        while (top > 0) {
            const Node& nd = nodes[stack[--top]];
            if (found == k && box_dist2(nd, x, y) > bd[k - 1]) {
                continue;
            }
            if (nd.left < 0) {
                for (int q = nd.lo; q < nd.hi; ++q) {
                    const int s = idx[q];
                    const int64_t dx = P[2 * s] - x, dy = P[2 * s + 1] - y;
                    const int64_t d2 = dx * dx + dy * dy;
                    if (found == k && (d2 > bd[k - 1] || (d2 == bd[k - 1] && s > bi[k - 1]))) {
                        continue;
                    }
                    int pos = found < k ? found : k - 1;
                    while (pos > 0 && (bd[pos - 1] > d2 || (bd[pos - 1] == d2 && bi[pos - 1] > s))) {
                        bd[pos] = bd[pos - 1];
                        bi[pos] = bi[pos - 1];
                        --pos;
                    }
                    bd[pos] = d2;
                    bi[pos] = s;
                    if (found < k) {
                        ++found;
                    }
                }
                continue;
            }
            /* nearer child last, so it is popped first */
            const bool left_first = box_dist2(nodes[nd.left], x, y) <= box_dist2(nodes[nd.right], x, y);
            stack[top++] = left_first ? nd.right : nd.left;
            stack[top++] = left_first ? nd.left : nd.right;
        }
        return found;
    }

private:
    static constexpr int kLeaf = 8;
    struct Node {
        int lo, hi;          /* range of idx */
        int left, right;     /* children, -1 for a leaf */
        int x0, y0, x1, y1;  /* bounding box */
    };

    static int64_t box_dist2(const Node& n, int x, int y)
    {
        const int64_t dx = x < n.x0 ? n.x0 - x : (x > n.x1 ? x - n.x1 : 0);
        const int64_t dy = y < n.y0 ? n.y0 - y : (y > n.y1 ? y - n.y1 : 0);
        return dx * dx + dy * dy;
    }

    int build(int lo, int hi)
    {
        const int id = (int)nodes.size();
        nodes.push_back(Node());
        Node nd;
        nd.lo = lo;
        nd.hi = hi;
        nd.left = nd.right = -1;
        nd.x0 = nd.y0 = INT32_MAX;
        nd.x1 = nd.y1 = INT32_MIN;
        for (int q = lo; q < hi; ++q) {
            const int s = idx[q];
            nd.x0 = std::min(nd.x0, P[2 * s]);
            nd.x1 = std::max(nd.x1, P[2 * s]);
            nd.y0 = std::min(nd.y0, P[2 * s + 1]);
            nd.y1 = std::max(nd.y1, P[2 * s + 1]);
        }
        if (hi - lo > kLeaf) {
            const int dim = (nd.x1 - nd.x0) >= (nd.y1 - nd.y0) ? 0 : 1;
            const int mid = lo + (hi - lo) / 2;
            std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
                const int va = P[2 * a + dim], vb = P[2 * b + dim];
                return va < vb || (va == vb && a < b);
            });
            nd.left = build(lo, mid);
            nd.right = build(mid, hi);
        }
        nodes[id] = nd;
        return id;
    }

    const int32_t* P;
    std::vector<int> idx;
    std::vector<Node> nodes;
};

/* Get3DPointonRefCam_factor (HPM.cpp:236-245) */
inline void point_ref_factor(const hpm_camera& c, int x, int y, float depth, float factor, float X[3])
{
    X[0] = depth * ((float)x - c.K[2] * factor) / (c.K[0] * factor);
    X[1] = depth * ((float)y - c.K[5] * factor) / (c.K[4] * factor);
    X[2] = depth;
}

} // namespace

extern "C" int hpm_support_points(const float* cost, int width, int height, int32_t* points_xy, int max_points)
{
    const int step = 5;
    int n = 0;
    for (int col = 0; col < width; col += step) {
        for (int row = 0; row < height; row += step) {
            float min_cost = 2.0f;
            int bx = -1, by = -1;
            const int cb = std::min(width, col + step), rb = std::min(height, row + step);
            for (int c = col; c < cb; ++c) {
                for (int r = row; r < rb; ++r) {
                    const float v = cost[(size_t)r * width + c];
                    if (v < 2.0f && min_cost > v) {
                        bx = c;
                        by = r;
                        min_cost = v;
                    }
                }
            }
            if (min_cost < 0.1f) {
                if (n < max_points && points_xy) {
                    points_xy[2 * n] = bx;
                    points_xy[2 * n + 1] = by;
                }
                ++n;
            }
        }
    }
    return n;
}

extern "C" int hpm_delaunay(const int32_t* points_xy, int n, int32_t* tris, int max_tris)
{
    if (n < 0 || (n > 0 && !points_xy)) {
        return -1;
    }
    if (n < 3) {
        return 0;
    }
    std::vector<hpm_dt::P2> pts((size_t)n);
    for (int i = 0; i < n; ++i) {
        pts[(size_t)i] = {points_xy[2 * i], points_xy[2 * i + 1]};
    }
    /* of points with equal coordinates the first one is triangulated */
    std::vector<int> byxy((size_t)n);
    for (int i = 0; i < n; ++i) {
        byxy[(size_t)i] = i;
    }
    std::sort(byxy.begin(), byxy.end(), [&](int a, int b) {
        const hpm_dt::P2 p = pts[(size_t)a], q = pts[(size_t)b];
        return p.x < q.x || (p.x == q.x && (p.y < q.y || (p.y == q.y && a < b)));
    });
    std::vector<char> dup((size_t)n, 0);
    for (size_t k = 1; k < byxy.size(); ++k) {
        const hpm_dt::P2 p = pts[(size_t)byxy[k - 1]], q = pts[(size_t)byxy[k]];
        if (p.x == q.x && p.y == q.y) {
            dup[(size_t)byxy[k]] = 1;
        }
    }
    std::vector<int> keep;
    for (int i = 0; i < n; ++i) {
        if (!dup[(size_t)i]) {
            keep.push_back(i);
        }
    }
    if (keep.size() < 3) {
        return 0;
    }
    const HostGrid g(pts, keep);
    /* the stars are independent: count, place, then write them in parallel */
    const int nk = (int)keep.size();
    std::vector<int> count((size_t)nk + 1, 0);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (int k = 0; k < nk; ++k) {
        int c = 0;
        const int r = hpm_dt::star(g, keep[(size_t)k], 4 * n + 16, [&](int, int) { ++c; });
        count[(size_t)k + 1] = c;
        bad += r < 0;
    }
    if (bad) {
        return -2;
    }
    for (int k = 0; k < nk; ++k) {
        count[(size_t)k + 1] += count[(size_t)k];
    }
    const int nt = count[(size_t)nk];
    if (tris) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int k = 0; k < nk; ++k) {
            const int p = keep[(size_t)k];
            int at = count[(size_t)k];
            hpm_dt::star(g, p, 4 * n + 16, [&](int a, int b) {
                if (at < max_tris) {
                    tris[3 * at] = p;
                    tris[3 * at + 1] = a;
                    tris[3 * at + 2] = b;
                }
                ++at;
            });
        }
    }
    return nt;
}

extern "C" void hpm_rasterize_triangles(const int32_t* pts, const int32_t* tris, int ntris, float* mask, int width,
                                        int height)
{
    uint32_t idx = 0;
    for (int t = 0; t < ntris; ++t) {
        const int x1 = pts[2 * tris[3 * t]], y1 = pts[2 * tris[3 * t] + 1];
        const int x2 = pts[2 * tris[3 * t + 1]], y2 = pts[2 * tris[3 * t + 1] + 1];
        const int x3 = pts[2 * tris[3 * t + 2]], y3 = pts[2 * tris[3 * t + 2] + 1];
        /* imageRC.contains for the three vertices (main.cpp:97) */
        if (x1 < 0 || x1 >= width || y1 < 0 || y1 >= height || x2 < 0 || x2 >= width || y2 < 0 || y2 >= height ||
            x3 < 0 || x3 >= width || y3 < 0 || y3 >= height) {
            continue;
        }
        const float L01 = (float)sqrt(pow((double)(x1 - x2), 2) + pow((double)(y1 - y2), 2));
        const float L02 = (float)sqrt(pow((double)(x1 - x3), 2) + pow((double)(y1 - y3), 2));
        const float L12 = (float)sqrt(pow((double)(x2 - x3), 2) + pow((double)(y2 - y3), 2));
        const float max_edge = std::max(L01, std::max(L02, L12));
        const float step = (float)(1.0 / max_edge);
        for (float p = 0; p < 1.0; p += step) {
            for (float q = 0; (double)q < 1.0 - (double)p; q += step) {
                const int x = (int)((double)(p * (float)x1 + q * (float)x2) + (1.0 - (double)p - (double)q) * x3);
                const int y = (int)((double)(p * (float)y1 + q * (float)y2) + (1.0 - (double)p - (double)q) * y3);
                if (x >= 0 && x < width && y >= 0 && y < height) {
                    mask[(size_t)y * width + x] = (float)(idx + 1.0);
                }
            }
        }
        idx++;
    }
}

extern "C" void hpm_plane_from_points(const hpm_camera* ref, const int32_t p[6], const float* depth, int width,
                                      float factor, float out[4])
{
    float X[3][3];
    for (int i = 0; i < 3; ++i) {
        const int x = p[2 * i], y = p[2 * i + 1];
        point_ref_factor(*ref, x, y, depth[(size_t)y * width + x], factor, X[i]);
    }
    /* null vector of [X_i 1] (cv::SVD::solveZ): n = (X2 - X1) x (X3 - X1),
     * w = -n . X1, in double */
    const double a[3] = {(double)X[1][0] - X[0][0], (double)X[1][1] - X[0][1], (double)X[1][2] - X[0][2]};
    const double b[3] = {(double)X[2][0] - X[0][0], (double)X[2][1] - X[0][1], (double)X[2][2] - X[0][2]};
    const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double w = -(n[0] * X[0][0] + n[1] * X[0][1] + n[2] * X[0][2]);
    /* HPM.cpp:834-841: unit normal, sign chosen so that w >= 0 */
    double norm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (w < 0) {
        norm = -norm;
    }
    out[0] = (float)(n[0] / norm);
    out[1] = (float)(n[1] / norm);
    out[2] = (float)(n[2] / norm);
    out[3] = (float)(w / norm);
}

extern "C" float hpm_depth_from_plane_factor(const hpm_camera* ref, const float pl[4], int x, int y, float factor)
{
    const hpm_camera& c = *ref;
    return -pl[3] * (c.K[0] * factor) /
           (((float)x - c.K[2] * factor) * pl[0] + (c.K[0] / c.K[4]) * ((float)y - c.K[5] * factor) * pl[1] +
            c.K[0] * factor * pl[2]);
}

extern "C" int hpm_knn_supplement(const hpm_camera* ref, const int32_t* pts, int npoints, const float* depth,
                                  int width, int height, float factor, int k, float depth_min, float depth_max,
                                  float* mask, float* planes, int num_threads)
{
    if (k < 3 || k > 32 || npoints < 0 || !ref || !depth || !mask || !planes) {
        return -1;
    }
    if (npoints < 3) {
        return 0;
    }
    /* triple order of HPM.cpp:982-1084: all triples of the first five
     * neighbours in lexicographic order, then, for each later neighbour m,
     * the lexicographic pairs before it with m */
    std::vector<int> order;
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b)
            for (int c = b + 1; c < 5; ++c) {
                order.push_back(a);
                order.push_back(b);
                order.push_back(c);
            }
    for (int m = 5; m < k; ++m)
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b) {
                order.push_back(a);
                order.push_back(b);
                order.push_back(m);
            }
    const KdTree tree(pts, npoints);
    const int nt = num_threads > 0 ? num_threads : 1;
    const int kk = std::min(k, npoints);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int i = 0; i < height; i++) {
        int64_t bd[32];
        int bi[32];
        for (int j = 0; j < width; j++) {
            if (mask[(size_t)i * width + j] > 0) {
                continue;
            }
            const int found = tree.knn(j, i, kk, bd, bi);
            if (found < 3) {
                continue;
            }
            int tri[6];
            bool ok = false;
            for (size_t t = 0; t + 2 < order.size(); t += 3) {
                if (order[t + 2] >= found) {
                    continue;
                }
                const int a = bi[order[t]], b = bi[order[t + 1]], c = bi[order[t + 2]];
                const int64_t cr = (int64_t)(pts[2 * b] - pts[2 * a]) * (pts[2 * c + 1] - pts[2 * a + 1]) -
                                   (int64_t)(pts[2 * b + 1] - pts[2 * a + 1]) * (pts[2 * c] - pts[2 * a]);
                if (cr != 0) {
                    tri[0] = pts[2 * a];
                    tri[1] = pts[2 * a + 1];
                    tri[2] = pts[2 * b];
                    tri[3] = pts[2 * b + 1];
                    tri[4] = pts[2 * c];
                    tri[5] = pts[2 * c + 1];
                    ok = true;
                    break;
                }
            }
            if (!ok) {
                continue;
            }
            float n4[4];
            hpm_plane_from_points(ref, tri, depth, width, factor, n4);
            const float d = hpm_depth_from_plane_factor(ref, n4, j, i, factor);
            if (d <= depth_max && d >= depth_min) {
                const size_t center = (size_t)i * width + j;
                mask[center] = -233.0f;
                memcpy(&planes[4 * center], n4, 16);
            }
        }
    }
    return 0;
}

extern "C" int hpm_multiscale_count(int width, int height)
{
    int max_size = std::max(width, height);
    int k = 0;
    while (max_size > 1000) {
        max_size /= 2;
        k++;
    }
    return k;
}
