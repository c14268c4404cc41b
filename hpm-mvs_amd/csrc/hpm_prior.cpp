/*
 * hpm_prior.cpp — host construction of the planar prior (include/hpm_prior.h).
 */
#include "hpm_prior.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

typedef __int128 i128;

struct Pt {
    int64_t x, y;
};

/* > 0 when c lies to the left of a->b (counter-clockwise in a y-up frame) */
inline int64_t orient(const Pt& a, const Pt& b, const Pt& c)
{
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}

/* > 0 when d lies strictly inside the circumcircle of the positively
 * oriented triangle (a, b, c); exact in 64 bits while every coordinate
 * difference is below 2^14 (12 D^4 < 2^63) */
inline int incircle64(const Pt& a, const Pt& b, const Pt& c, const Pt& d)
{
    const int64_t adx = a.x - d.x, ady = a.y - d.y;
    const int64_t bdx = b.x - d.x, bdy = b.y - d.y;
    const int64_t cdx = c.x - d.x, cdy = c.y - d.y;
    const int64_t ad = adx * adx + ady * ady;
    const int64_t bd = bdx * bdx + bdy * bdy;
    const int64_t cd = cdx * cdx + cdy * cdy;
    const int64_t det = adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
    return det > 0 ? 1 : (det < 0 ? -1 : 0);
}

/* the same with 128-bit products, for triangles of the enclosing triangle */
inline int incircle(const Pt& a, const Pt& b, const Pt& c, const Pt& d)
{
    const i128 adx = a.x - d.x, ady = a.y - d.y;
    const i128 bdx = b.x - d.x, bdy = b.y - d.y;
    const i128 cdx = c.x - d.x, cdy = c.y - d.y;
    const i128 ad = adx * adx + ady * ady;
    const i128 bd = bdx * bdx + bdy * bdy;
    const i128 cd = cdx * cdx + cdy * cdy;
    const i128 det = adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
    return det > 0 ? 1 : (det < 0 ? -1 : 0);
}

struct Tri {
    int v[3];   /* positively oriented */
    int nb[3];  /* neighbour across the edge opposite v[i], -1 if none */
    bool alive;
};

/* Bowyer-Watson insertion with a visibility walk for point location. */
class Delaunay {
public:
    explicit Delaunay(const std::vector<Pt>& pts) : P(pts)
    {
        int64_t maxc = 1;
        for (const Pt& p : P) {
            const int64_t ax = p.x < 0 ? -p.x : p.x, ay = p.y < 0 ? -p.y : p.y;
            maxc = std::max(maxc, std::max(ax, ay));
        }
        const int64_t S = 64 * (maxc + 16);
        n_real = (int)P.size();
        small_coords = 2 * maxc < ((int64_t)1 << 14);
        P.push_back({-S, -S});
        P.push_back({3 * S, -S});
        P.push_back({-S, 3 * S});
        T.reserve(2 * (size_t)n_real + 16);
        marks.assign(2 * (size_t)n_real + 16, 0);
        Tri t;
        t.v[0] = n_real;
        t.v[1] = n_real + 1;
        t.v[2] = n_real + 2;
        t.nb[0] = t.nb[1] = t.nb[2] = -1;
        t.alive = true;
        T.push_back(t);
        last = 0;
    }

    bool insert(int pi)
    {
        const Pt& p = P[pi];
        int t = locate(p);
        if (t < 0) {
            return false;
        }
        for (int i = 0; i < 3; ++i) {
            const Pt& q = P[T[t].v[i]];
            if (q.x == p.x && q.y == p.y) {
                return false; /* duplicate point */
            }
        }
        /* cavity: triangles whose circumcircle strictly contains p */
        cavity.clear();
        stack.clear();
        mark_gen++;
        if (marks.size() < T.size()) {
            marks.resize(T.size() * 2, 0);
        }
        stack.push_back(t);
        marks[t] = mark_gen;
        while (!stack.empty()) {
            const int c = stack.back();
            stack.pop_back();
            cavity.push_back(c);
            for (int e = 0; e < 3; ++e) {
                const int n = T[c].nb[e];
                if (n < 0 || marks[n] == mark_gen) {
                    continue;
                }
                const Tri& tn = T[n];
                const bool real = tn.v[0] < n_real && tn.v[1] < n_real && tn.v[2] < n_real && small_coords;
                const int in = real ? incircle64(P[tn.v[0]], P[tn.v[1]], P[tn.v[2]], p)
                                    : incircle(P[tn.v[0]], P[tn.v[1]], P[tn.v[2]], p);
                if (in > 0) {
                    marks[n] = mark_gen;
                    stack.push_back(n);
                }
            }
        }
        /* boundary edges (a, b) of the cavity with the outside neighbour */
        boundary.clear();
        for (int c : cavity) {
            for (int e = 0; e < 3; ++e) {
                const int n = T[c].nb[e];
                if (n >= 0 && marks[n] == mark_gen) {
                    continue;
                }
                Edge ed;
                ed.a = T[c].v[(e + 1) % 3];
                ed.b = T[c].v[(e + 2) % 3];
                ed.outside = n;
                ed.outside_edge = -1;
                if (n >= 0) {
                    for (int k = 0; k < 3; ++k) {
                        if (T[n].nb[k] == c) {
                            ed.outside_edge = k;
                        }
                    }
                }
                boundary.push_back(ed);
            }
        }
        for (int c : cavity) {
            T[c].alive = false;
            free_slots.push_back(c);
        }
        /* new triangles (a, b, p), linked around p through the map a -> tri */
        created.clear();
        for (const Edge& ed : boundary) {
            int slot;
            if (!free_slots.empty()) {
                slot = free_slots.back();
                free_slots.pop_back();
            } else {
                slot = (int)T.size();
                T.push_back(Tri());
            }
            Tri& nt = T[slot];
            nt.v[0] = ed.a;
            nt.v[1] = ed.b;
            nt.v[2] = pi;
            nt.nb[2] = ed.outside;
            nt.nb[0] = -1;
            nt.nb[1] = -1;
            nt.alive = true;
            if (ed.outside >= 0 && ed.outside_edge >= 0) {
                T[ed.outside].nb[ed.outside_edge] = slot;
            }
            created.push_back(slot);
        }
        /* edge (b, p) of triangle (a, b, p) is opposite a (nb[0]); it is
         * shared with the triangle whose first vertex is b */
        first_of.clear();
        for (int s : created) {
            first_of.emplace_back(T[s].v[0], s);
        }
        std::sort(first_of.begin(), first_of.end());
        for (int s : created) {
            const int b = T[s].v[1];
            auto it = std::lower_bound(first_of.begin(), first_of.end(), std::make_pair(b, -1));
            if (it != first_of.end() && it->first == b) {
                const int o = it->second;
                T[s].nb[0] = o;
                T[o].nb[1] = s;
            }
        }
        if (marks.size() < T.size()) {
            marks.resize(T.size() * 2, 0);
        }
        last = created.empty() ? last : created[0];
        return true;
    }

    /* triangles made only of real points, in slot order */
    void output(std::vector<int32_t>& out) const
    {
        out.reserve(3 * T.size());
        for (const Tri& t : T) {
            if (!t.alive || t.v[0] >= n_real || t.v[1] >= n_real || t.v[2] >= n_real) {
                continue;
            }
            out.push_back(t.v[0]);
            out.push_back(t.v[1]);
            out.push_back(t.v[2]);
        }
    }

private:
    struct Edge {
        int a, b, outside, outside_edge;
    };

    int locate(const Pt& p)
    {
        int t = last;
        if (t < 0 || t >= (int)T.size() || !T[t].alive) {
            t = -1;
            for (int i = (int)T.size() - 1; i >= 0; --i) {
                if (T[i].alive) {
                    t = i;
                    break;
                }
            }
        }
        const size_t limit = 4 * T.size() + 64;
        for (size_t step = 0; step < limit && t >= 0; ++step) {
            int next = -1;
            for (int e = 0; e < 3; ++e) {
                const Pt& a = P[T[t].v[(e + 1) % 3]];
                const Pt& b = P[T[t].v[(e + 2) % 3]];
                if (orient(a, b, p) < 0) {
                    next = T[t].nb[e];
                    break;
                }
            }
            if (next < 0) {
                return t; /* inside or on the boundary of t */
            }
            t = next;
        }
        /* fall back to a scan (never expected on a Delaunay triangulation) */
        for (int i = 0; i < (int)T.size(); ++i) {
            if (!T[i].alive) {
                continue;
            }
            bool in = true;
            for (int e = 0; e < 3 && in; ++e) {
                in = orient(P[T[i].v[(e + 1) % 3]], P[T[i].v[(e + 2) % 3]], p) >= 0;
            }
            if (in) {
                return i;
            }
        }
        return -1;
    }

    std::vector<Pt> P;
    std::vector<Tri> T;
    int n_real;
    bool small_coords; /* real points differ by less than 2^14 per axis */
    int last;
    std::vector<int> cavity, stack, free_slots, created;
    std::vector<Edge> boundary;
    std::vector<std::pair<int, int>> first_of;
    std::vector<uint32_t> marks;
    uint32_t mark_gen = 0;
};

/* position of (x, y) along a Hilbert curve over [0, 2^16)^2 */
inline uint64_t hilbert_key(uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

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
    std::vector<Pt> pts((size_t)n);
    for (int i = 0; i < n; ++i) {
        pts[i] = {points_xy[2 * i], points_xy[2 * i + 1]};
    }
    Delaunay dt(pts);
    /* insertion along a Hilbert curve: short walks and small cavities (in
     * grid order the cavities of near-lattice support points are large) */
    int32_t minx = (int32_t)pts[0].x, miny = (int32_t)pts[0].y;
    for (const Pt& p : pts) {
        minx = (int32_t)std::min<int64_t>(minx, p.x);
        miny = (int32_t)std::min<int64_t>(miny, p.y);
    }
    std::vector<std::pair<uint64_t, int>> order((size_t)n);
    for (int i = 0; i < n; ++i) {
        order[i] = {hilbert_key((uint32_t)(pts[i].x - minx) & 0xffffu, (uint32_t)(pts[i].y - miny) & 0xffffu), i};
    }
    std::sort(order.begin(), order.end());
    for (const auto& o : order) {
        dt.insert(o.second);
    }
    std::vector<int32_t> out;
    dt.output(out);
    const int nt = (int)(out.size() / 3);
    if (tris) {
        memcpy(tris, out.data(), sizeof(int32_t) * 3 * (size_t)std::min(nt, max_tris));
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
