/*
 * hpm_dt.h — the Delaunay triangulation of the planar prior's support points
 * (the reference's cv::Subdiv2D triangulation, HPM.cpp:787-809), one
 * algorithm for the host library and the device kernels.
 *
 * Every point p finds its own star, the triangles around it: it starts from
 * its nearest neighbour q0 (always a Delaunay edge: the circle on p q0 as
 * diameter holds no other point), then turns counter-clockwise: the third
 * vertex on the left of an edge p->q is the left point whose circle through
 * p and q holds no other left point. Candidates come from a uniform grid of
 * cells in growing square rings around p's cell; the search stops once the
 * best circle's diameter fits inside the rings scanned (every point in the
 * circle lies within a diameter of p), so the result does not depend on the
 * grid or the scan order, only on the predicates:
 *   - orient: exact in 64 bits for coordinates below 2^30;
 *   - incircle: exact in 64 bits for coordinate differences below 2^14;
 *     co-circular ties are broken by lifting point i to x^2 + y^2 + eps_i
 *     with eps_0 >> eps_1 >> ... (a simulation of simplicity on the lifted
 *     coordinate): the sign is that of the first non-zero cofactor of the
 *     lifted column, in ascending point index. The triangulation is then
 *     unique, so every correct algorithm with these predicates yields it.
 * A point emits the triangles of its star whose smallest vertex index is
 * its own, as (p, a, b) counter-clockwise, in star order: counter-clockwise
 * from q0 when the star closes, else counter-clockwise from the hull edge
 * on its clockwise side. Points in ascending index order then give the
 * canonical triangle list.
 */
#ifndef HPM_DT_H
#define HPM_DT_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define HPM_DT_FN __host__ __device__ inline
/* one copy of the larger predicates: divergent lanes of a wave run
 * different parts of the walk, and inlined copies would overflow the
 * instruction cache */
#define HPM_DT_CALL __host__ __device__ __attribute__((noinline))
#else
#define HPM_DT_FN inline
#define HPM_DT_CALL inline
#endif

namespace hpm_dt {

struct P2 {
    int32_t x, y;
};

/* > 0 when c lies left of a->b (counter-clockwise in a y-up frame) */
HPM_DT_FN int64_t orient(P2 a, P2 b, P2 c)
{
    return (int64_t)(b.x - a.x) * (int64_t)(c.y - a.y) - (int64_t)(b.y - a.y) * (int64_t)(c.x - a.x);
}

HPM_DT_FN int sgn64(int64_t v)
{
    return (v > 0) - (v < 0);
}

/* the incircle determinant of (a, b, c, d), > 0 when d is inside the circle
 * of the counter-clockwise triangle (a, b, c); exact while every
 * coordinate difference is below 2^14 (12 D^4 < 2^63) */
HPM_DT_FN int64_t incircle_det(P2 a, P2 b, P2 c, P2 d)
{
    const int64_t adx = a.x - d.x, ady = a.y - d.y;
    const int64_t bdx = b.x - d.x, bdy = b.y - d.y;
    const int64_t cdx = c.x - d.x, cdy = c.y - d.y;
    const int64_t ad = adx * adx + ady * ady;
    const int64_t bd = bdx * bdx + bdy * bdy;
    const int64_t cd = cdx * cdx + cdy * cdy;
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}

/* incircle with the ties broken as described above: the determinant of
 * rows (x, y, x^2 + y^2 + eps_i, 1) in the order a, b, c, d; the cofactors
 * of the lifted column are +O(b,c,d), -O(a,c,d), +O(a,b,d), -O(a,b,c).
 * Non-zero whenever (a, b, c) is a proper triangle. */
HPM_DT_CALL int incircle(P2 a, int ia, P2 b, int ib, P2 c, int ic, P2 d, int id)
{
    const int64_t det = incircle_det(a, b, c, d);
    if (det != 0) {
        return sgn64(det);
    }
    const int idx[4] = {ia, ib, ic, id};
    const int64_t cof[4] = {orient(b, c, d), -orient(a, c, d), orient(a, b, d), -orient(a, b, c)};
    int done = 0; /* bit k: cofactor k consumed */
    for (int round = 0; round < 4; ++round) {
        int m = -1;
        for (int k = 0; k < 4; ++k) {
            if (!((done >> k) & 1) && (m < 0 || idx[k] < idx[m])) {
                m = k;
            }
        }
        done |= 1 << m;
        if (cof[m] != 0) {
            return sgn64(cof[m]);
        }
    }
    return 0;
}

/* Grid interface (template parameter G):
 *   P2 pt(int i) const;              the point i
 *   int gw, gh, cs;                  cells across, down, and the cell size
 *   int cx(P2 p) const, cy(P2 p) const;   the cell of a point
 *   int cxd(double x) const, cyd(double y) const;   the cell column / row of
 *       a coordinate, clamped to [-1, gw] / [-1, gh]
 *   double rowy(int y) const, colx(int x) const;   the first pixel row /
 *       column of cell row y / column x
 *   template <class F> void visit_row(int y, int x0, int x1, F f) const;
 *   template <class F> void visit_col(int x, int y0, int y1, F f) const;
 *       f(i) for the points of the cells (x0..x1, y) / (x, y0..y1), ranges
 *       inside the grid; a grid may skip empty cells in one step
 *   int nhull; P2 hull(int k) const; the convex hull's vertices
 * A point lies at most cs - 1 from its cell's corner in each axis. */

/* nothing strictly left of a->b: every hull vertex is on the right or on
 * the line */
template <class G>
HPM_DT_FN bool nothing_left(const G& g, P2 a, P2 b)
{
    for (int k = 0; k < g.nhull; ++k) {
        if (orient(a, b, g.hull(k)) > 0) {
            return false;
        }
    }
    return true;
}

/* visits the cells of square ring r around (cx, cy) inside the grid;
 * returns false when the ring lies wholly outside the grid */
template <class G, class F>
HPM_DT_FN bool visit_ring(const G& g, int cx, int cy, int r, F&& f)
{
    const int x0 = cx - r, x1 = cx + r, y0 = cy - r, y1 = cy + r;
    if (x0 < 0 && y0 < 0 && x1 >= g.gw && y1 >= g.gh) {
        return false;
    }
    if (r == 0) {
        g.visit_row(cy, cx, cx, f);
        return true;
    }
    const int xa = x0 < 0 ? 0 : x0, xb = x1 >= g.gw ? g.gw - 1 : x1;
    if (y0 >= 0) {
        g.visit_row(y0, xa, xb, f);
    }
    if (y1 < g.gh) {
        g.visit_row(y1, xa, xb, f);
    }
    const int ya = y0 + 1 < 0 ? 0 : y0 + 1, yb = y1 - 1 >= g.gh ? g.gh - 1 : y1 - 1;
    if (ya <= yb) {
        if (x0 >= 0) {
            g.visit_col(x0, ya, yb, f);
        }
        if (x1 < g.gw) {
            g.visit_col(x1, ya, yb, f);
        }
    }
    return true;
}

/* the nearest other point of p (ties to the smaller index), -1 if none */
template <class G>
HPM_DT_CALL int nearest(const G& g, int p)
{
    const P2 pp = g.pt(p);
    const int cx = g.cx(pp), cy = g.cy(pp);
    int best = -1;
    int64_t bd = 0;
    for (int r = 0;; ++r) {
        const bool any = visit_ring(g, cx, cy, r, [&](int i) {
            if (i == p) {
                return;
            }
            const P2 q = g.pt(i);
            const int64_t dx = q.x - pp.x, dy = q.y - pp.y;
            const int64_t d = dx * dx + dy * dy;
            if (best < 0 || d < bd || (d == bd && i < best)) {
                best = i;
                bd = d;
            }
        });
        if (!any) {
            return best;
        }
        /* unscanned points lie more than r * cs from p */
        const int64_t reach = (int64_t)r * g.cs;
        if (best >= 0 && bd <= reach * reach) {
            return best;
        }
    }
}

/* p lies on the boundary of the convex hull (on a hull edge) */
template <class G>
HPM_DT_FN bool on_hull(const G& g, P2 p)
{
    for (int k = 0; k < g.nhull; ++k) {
        const P2 a = g.hull(k), b = g.hull(k + 1 < g.nhull ? k + 1 : 0);
        if (orient(a, b, p) == 0 && (p.x - a.x) * (int64_t)(p.x - b.x) <= 0 && (p.y - a.y) * (int64_t)(p.y - b.y) <= 0) {
            return true;
        }
    }
    return g.nhull < 3;
}

/* the third vertex of the Delaunay triangle on the left of a->b (a is the
 * point whose rings are scanned), -1 when nothing lies on the left. Some
 * point lies strictly left whenever a or b lies inside the hull (a line
 * through an interior point has hull vertices on both sides); boundary
 * says whether both lie on the hull's boundary. */
template <class G>
HPM_DT_CALL int third(const G& g, int a, int b, bool boundary)
{
    const P2 pa = g.pt(a), pb = g.pt(b);
    if (boundary && nothing_left(g, pa, pb)) {
        return -1;
    }
    const int cx = g.cx(pa), cy = g.cy(pa);
    int best = -1;
    auto consider = [&](int i) {
        if (i == a || i == b || i == best) {
            return;
        }
        const P2 d = g.pt(i);
        if (orient(pa, pb, d) <= 0) {
            return;
        }
        if (best < 0 || incircle(pa, a, pb, b, g.pt(best), best, d, i) > 0) {
            best = i;
        }
    };
    const double abx = (double)pb.x - pa.x, aby = (double)pb.y - pa.y;
    /* the cells of row y that can hold points left of a->b, clipped to
     * [xs, xe]; false when none */
    auto left_row = [&](int y, int* xs, int* xe) {
        const double y0 = g.rowy(y), y1 = y0 + (g.cs - 1);
        double lo = -1e300, hi = 1e300;
        if (aby > 0.0) {
            hi = pa.x + abx * ((abx >= 0.0 ? y1 : y0) - pa.y) / aby + 1e-6;
        } else if (aby < 0.0) {
            lo = pa.x + abx * ((abx >= 0.0 ? y1 : y0) - pa.y) / aby - 1e-6;
        } else if ((abx > 0.0 && !(y1 > pa.y)) || (abx < 0.0 && !(y0 < pa.y))) {
            return false;
        }
        const int a0 = g.cxd(lo), a1 = g.cxd(hi);
        *xs = a0 > *xs ? a0 : *xs;
        *xe = a1 < *xe ? a1 : *xe;
        return *xs <= *xe;
    };
    /* the same for the rows of column x (a column band of cs pixels) */
    auto left_col = [&](int x, int* ys, int* ye) {
        const double x0 = g.colx(x), x1 = x0 + (g.cs - 1);
        double lo = -1e300, hi = 1e300;
        /* (b - a) x (p - a) > 0: abx (y - a.y) > aby (x - a.x) */
        if (abx > 0.0) {
            lo = pa.y + aby * ((aby >= 0.0 ? x0 : x1) - pa.x) / abx - 1e-6;
        } else if (abx < 0.0) {
            hi = pa.y + aby * ((aby >= 0.0 ? x0 : x1) - pa.x) / abx + 1e-6;
        } else if ((aby > 0.0 && !(x0 < pa.x)) || (aby < 0.0 && !(x1 > pa.x))) {
            return false;
        }
        const int a0 = g.cyd(lo), a1 = g.cyd(hi);
        *ys = a0 > *ys ? a0 : *ys;
        *ye = a1 < *ye ? a1 : *ye;
        return *ys <= *ye;
    };
    /* a first candidate from the rings around a (at least rings 0 and 1),
     * each side of a ring clipped to the left half-plane */
    for (int r = 0;; ++r) {
        const int x0 = cx - r, x1 = cx + r, y0 = cy - r, y1 = cy + r;
        if (x0 < 0 && y0 < 0 && x1 >= g.gw && y1 >= g.gh) {
            return best;
        }
        const int xa = x0 < 0 ? 0 : x0, xb = x1 >= g.gw ? g.gw - 1 : x1;
        for (int side = 0; side < 2; ++side) {
            const int y = side == 0 ? y0 : y1;
            if (y < 0 || y >= g.gh || (side == 1 && r == 0)) {
                continue;
            }
            int xs = xa, xe = xb;
            if (left_row(y, &xs, &xe)) {
                g.visit_row(y, xs, xe, consider);
            }
        }
        const int ya = y0 + 1 < 0 ? 0 : y0 + 1, yb = y1 - 1 >= g.gh ? g.gh - 1 : y1 - 1;
        for (int side = 0; side < 2 && ya <= yb; ++side) {
            const int x = side == 0 ? x0 : x1;
            if (x < 0 || x >= g.gw) {
                continue;
            }
            int ys = ya, ye = yb;
            if (left_col(x, &ys, &ye)) {
                g.visit_col(x, ys, ye, consider);
            }
        }
        if (best >= 0 && r >= 1) {
            break;
        }
    }
    /* On the left of a->b the circles through a and b are nested, so the
     * left part of the final circle lies inside the left part of the
     * current one, and when the current centre lies right of a->b that
     * part (a minor segment: every point sees a b at 90 degrees or more)
     * lies inside the disk on a b as diameter. The scan covers that
     * region row by row, outwards from a's row, each row over the cells its
     * chord of the region and the left half-plane reach (bounds widened for
     * the roundings); the region is recomputed as better candidates shrink
     * it, so every row scanned, and every row left out, is judged against
     * a region that contains the final one. */
    const double lab = sqrt(abx * abx + aby * aby);
    double ux = 0.0, uy = 0.0, rad = 0.0;
    int cur = -1, gy0 = 0, gy1 = -1;
    auto region = [&]() {
        cur = best;
        const P2 pc = g.pt(best);
        const double qx = (double)pc.x - pa.x, qy = (double)pc.y - pa.y;
        const double den = 2.0 * (double)orient(pa, pb, pc);
        const double b2 = abx * abx + aby * aby, q2 = qx * qx + qy * qy;
        const double cx = (qy * b2 - aby * q2) / den, cy = (abx * q2 - qx * b2) / den;
        /* signed distance of the centre from the line, > 0 on the left */
        const double side = (abx * cy - aby * cx) / lab;
        if (side < -1e-6 * (lab + 1.0)) {
            ux = pa.x + 0.5 * abx;
            uy = pa.y + 0.5 * aby;
            rad = 0.5 * lab * (1.0 + 1e-9) + 1e-6;
        } else {
            ux = pa.x + cx;
            uy = pa.y + cy;
            rad = sqrt(cx * cx + cy * cy) * (1.0 + 1e-9) + 1e-6;
        }
        gy0 = g.cyd(uy - rad);
        gy1 = g.cyd(uy + rad);
    };
    auto scan_row = [&](int y) {
        if (best != cur) {
            region();
        }
        if (y < gy0 || y > gy1 || y < 0 || y >= g.gh) {
            return;
        }
        const double y0 = g.rowy(y), y1 = y0 + (g.cs - 1);
        const double dyc = uy < y0 ? y0 - uy : (uy > y1 ? uy - y1 : 0.0);
        if (dyc > rad) {
            return;
        }
        const double hw = sqrt(rad * rad - dyc * dyc);
        double lo = ux - hw, hi = ux + hw;
        /* the left half-plane within the row band: (b - a) x (p - a) > 0 */
        if (aby > 0.0) {
            hi = fmin(hi, pa.x + abx * ((abx >= 0.0 ? y1 : y0) - pa.y) / aby + 1e-6);
        } else if (aby < 0.0) {
            lo = fmax(lo, pa.x + abx * ((abx >= 0.0 ? y1 : y0) - pa.y) / aby - 1e-6);
        } else if ((abx > 0.0 && !(y1 > pa.y)) || (abx < 0.0 && !(y0 < pa.y))) {
            return;
        }
        const int xa = g.cxd(lo), xb = g.cxd(hi);
        const int xs = xa < 0 ? 0 : xa, xe = xb >= g.gw ? g.gw - 1 : xb;
        if (xs <= xe) {
            g.visit_row(y, xs, xe, consider);
        }
    };
    region();
    const int ya = cy;
    for (int d = 0;; ++d) {
        if (ya - d < gy0 && ya + d > gy1) {
            break;
        }
        scan_row(ya - d);
        if (d > 0) {
            scan_row(ya + d);
        }
    }
    return best;
}

/* The star of p: calls emit(a, b) for each triangle (p, a, b) of the star,
 * counter-clockwise, whose smallest vertex index is p, in star order.
 * Returns the number of triangles of the star, or -1 after max_steps. */
template <class G, class E>
HPM_DT_FN int star(const G& g, int p, int max_steps, E&& emit)
{
    const int q0 = nearest(g, p);
    if (q0 < 0) {
        return 0;
    }
    const bool pb = on_hull(g, g.pt(p));
    /* the counter-clockwise walk from q0; the first kBuf emitted pairs are
     * kept, so a closed star need not be walked again */
    const int kBuf = 16;
    int buf[2 * kBuf];
    int nbuf = 0, ne = 0;
    int q = q0, steps = 0;
    bool closed = false;
    for (;;) {
        const int c = third(g, p, q, pb && on_hull(g, g.pt(q)));
        if (c < 0) {
            break;
        }
        if (p < q && p < c) {
            if (nbuf < kBuf) {
                buf[2 * nbuf] = q;
                buf[2 * nbuf + 1] = c;
                ++nbuf;
            }
            ++ne;
        }
        q = c;
        if (q == q0) {
            closed = true;
            break;
        }
        if (++steps > max_steps) {
            return -1;
        }
    }
    int start = q0;
    if (closed && ne == nbuf) {
        for (int k = 0; k < nbuf; ++k) {
            emit(buf[2 * k], buf[2 * k + 1]);
        }
        return steps + 1;
    }
    if (!closed) {
        /* a star that does not close starts at its clockwise hull edge:
         * clockwise from q0, the triangle (q, p, c) with c left of q->p */
        q = q0;
        steps = 0;
        for (;;) {
            const int c = third(g, q, p, pb && on_hull(g, g.pt(q)));
            if (c < 0) {
                break;
            }
            q = c;
            if (++steps > max_steps) {
                return -1;
            }
        }
        start = q;
    }
    int n = 0;
    q = start;
    steps = 0;
    for (;;) {
        const int c = third(g, p, q, pb && on_hull(g, g.pt(q)));
        if (c < 0) {
            break;
        }
        if (p < q && p < c) {
            emit(q, c);
        }
        ++n;
        q = c;
        if (q == start || ++steps > max_steps) {
            break;
        }
    }
    return steps > max_steps ? -1 : n;
}

} // namespace hpm_dt

#endif
