/*
 * hpm_prior.hip — the planar prior's per-pixel and per-triangle stages on
 * the device (main.cpp:84-162, 238-314, 387-454 with HPM.cpp:759-1123):
 * scale maps, support-point blocks, triangle rasterisation and plane fits,
 * the KNN supplement and the prior assembly. They compute what the host
 * functions of hpm_prior.cpp and the assembly loops of hpm_pipeline.cpp
 * compute, in the same arithmetic (no contraction; float where those use
 * float, double where they use double), so the prior is the same bit for
 * bit. The Delaunay triangulation runs the star walk of hpm_dt.h, the
 * same algorithm as the host's hpm_delaunay, over the support-point blocks.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "hpm_device.h"
#include "hpm_dt.h"

namespace {

/* hpm_resize_bilinear_c (hpm_io.cpp) for one output texel of a strided map */
__device__ float resize_texel(const float* src, int stride, int sw, int sh, int dw, int dh, int x, int y)
{
    const double sx_scale = (double)sw / dw;
    const double sy_scale = (double)sh / dh;
    float fy = (float)((y + 0.5) * sy_scale - 0.5);
    int y0 = (int)floorf(fy);
    fy -= (float)y0;
    int y1 = y0 + 1;
    if (y0 < 0) {
        y0 = 0;
        y1 = 0;
        fy = 0.0f;
    }
    if (y1 >= sh) {
        y0 = sh - 1;
        y1 = sh - 1;
        fy = 0.0f;
    }
    float fx = (float)((x + 0.5) * sx_scale - 0.5);
    int x0 = (int)floorf(fx);
    fx -= (float)x0;
    int x1 = x0 + 1;
    if (x0 < 0) {
        x0 = 0;
        x1 = 0;
        fx = 0.0f;
    }
    if (x1 >= sw) {
        x0 = sw - 1;
        x1 = sw - 1;
        fx = 0.0f;
    }
    const float a = src[((size_t)y0 * sw + x0) * stride];
    const float b = src[((size_t)y0 * sw + x1) * stride];
    const float d = src[((size_t)y1 * sw + x0) * stride];
    const float e = src[((size_t)y1 * sw + x1) * stride];
    const float top = a * (1.0f - fx) + b * fx;
    const float bot = d * (1.0f - fx) + e * fx;
    return top * (1.0f - fy) + bot * fy;
}

/* Get3DPointonRefCam_factor (HPM.cpp:236-245) */
__device__ __forceinline__ void point_ref_factor(const hpm_camera& c, int x, int y, float depth, float factor,
                                                 float X[3])
{
    X[0] = depth * ((float)x - c.K[2] * factor) / (c.K[0] * factor);
    X[1] = depth * ((float)y - c.K[5] * factor) / (c.K[4] * factor);
    X[2] = depth;
}

/* GetPriorPlaneParams_factor (HPM.cpp:811-844) as hpm_plane_from_points:
 * the null vector of [X_i 1] in double, unit normal with w >= 0 */
__device__ float4 plane_from_points(const hpm_camera& ref, const int32_t p[6], const float* depth, int width,
                                    float factor)
{
    float X[3][3];
    for (int i = 0; i < 3; ++i) {
        const int x = p[2 * i], y = p[2 * i + 1];
        point_ref_factor(ref, x, y, depth[(size_t)y * width + x], factor, X[i]);
    }
    const double a0 = (double)X[1][0] - X[0][0], a1 = (double)X[1][1] - X[0][1], a2 = (double)X[1][2] - X[0][2];
    const double b0 = (double)X[2][0] - X[0][0], b1 = (double)X[2][1] - X[0][1], b2 = (double)X[2][2] - X[0][2];
    const double n0 = a1 * b2 - a2 * b1, n1 = a2 * b0 - a0 * b2, n2 = a0 * b1 - a1 * b0;
    const double w = -(n0 * X[0][0] + n1 * X[0][1] + n2 * X[0][2]);
    double norm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (w < 0) {
        norm = -norm;
    }
    return make_float4((float)(n0 / norm), (float)(n1 / norm), (float)(n2 / norm), (float)(w / norm));
}

/* GetDepthFromPlaneParam_factor (HPM.cpp:846-849), hpm_depth_from_plane_factor */
__device__ __forceinline__ float depth_from_plane_factor(const hpm_camera& c, float4 pl, int x, int y, float factor)
{
    return -pl.w * (c.K[0] * factor) /
           (((float)x - c.K[2] * factor) * pl.x + (c.K[0] / c.K[4]) * ((float)y - c.K[5] * factor) * pl.y +
            c.K[0] * factor * pl.z);
}

} // namespace

/* depth (the .w of the finished hypotheses) and cost at the prior's scale:
 * copied at 1x, resized with the host's bilinear resize otherwise */
__global__ __launch_bounds__(256) void hpm_k_prior_maps(const float* hyp4, const float* cost, int W, int H,
                                                        float* depth_s, float* cost_s, int ws, int hs)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= ws || y >= hs) {
        return;
    }
    const size_t q = (size_t)y * ws + x;
    if (ws == W && hs == H) {
        depth_s[q] = hyp4[4 * q + 3];
        cost_s[q] = cost[q];
    } else {
        depth_s[q] = resize_texel(hyp4 + 3, 4, W, H, ws, hs, x, y);
        cost_s[q] = resize_texel(cost, 1, W, H, ws, hs, x, y);
    }
}

/* GetSupportPoints (HPM.cpp:759-785) per 5x5 block, blocks numbered
 * column-major (the reference's scan order): the first strict minimum of
 * the costs below 2 in column-major order, kept if below 0.1; packed x |
 * y << 16, or -1 */
__global__ __launch_bounds__(256) void hpm_k_support_blocks(const float* cost, int w, int h, int32_t* cand)
{
    const int nbx = (w + 4) / 5, nby = (h + 4) / 5;
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nbx * nby) {
        return;
    }
    const int col = (b / nby) * 5, row = (b % nby) * 5;
    float min_cost = 2.0f;
    int bx = -1, by = -1;
    const int cb = min(w, col + 5), rb = min(h, row + 5);
    for (int c = col; c < cb; ++c) {
        for (int r = row; r < rb; ++r) {
            const float v = cost[(size_t)r * w + c];
            if (v < 2.0f && min_cost > v) {
                bx = c;
                by = r;
                min_cost = v;
            }
        }
    }
    cand[b] = min_cost < 0.1f ? (bx | (by << 16)) : -1;
}

/* the triangle rasterisation of main.cpp:96-116 (hpm_rasterize_triangles):
 * one wave per triangle, lane l taking the outer steps l, l + 64, ... (the
 * steps are the float sums p += step, advanced 64 at a time per lane, so
 * every lane sees the values of the sequential loop); later triangles
 * overwrite earlier ones there, so a pixel ends with the largest index + 1
 * that reaches it (atomicMax) */
__global__ __launch_bounds__(256) void hpm_k_raster(const int32_t* pts, const int32_t* tris, const int32_t* tri_idx,
                                                    int ntris, uint32_t* mask, int width, int height)
{
    const int t = (int)(blockIdx.x * 4 + threadIdx.x / 64);
    const int lane = (int)(threadIdx.x % 64);
    if (t >= ntris || tri_idx[t] < 0) {
        return;
    }
    const uint32_t val = (uint32_t)tri_idx[t] + 1u;
    const int x1 = pts[2 * tris[3 * t]], y1 = pts[2 * tris[3 * t] + 1];
    const int x2 = pts[2 * tris[3 * t + 1]], y2 = pts[2 * tris[3 * t + 1] + 1];
    const int x3 = pts[2 * tris[3 * t + 2]], y3 = pts[2 * tris[3 * t + 2] + 1];
    const double d01x = (double)(x1 - x2), d01y = (double)(y1 - y2);
    const double d02x = (double)(x1 - x3), d02y = (double)(y1 - y3);
    const double d12x = (double)(x2 - x3), d12y = (double)(y2 - y3);
    const float L01 = (float)sqrt(d01x * d01x + d01y * d01y);
    const float L02 = (float)sqrt(d02x * d02x + d02y * d02y);
    const float L12 = (float)sqrt(d12x * d12x + d12y * d12y);
    const float max_edge = fmaxf(L01, fmaxf(L02, L12));
    const float step = (float)(1.0 / max_edge);
    float p = 0;
    for (int k = 0; k < lane; ++k) {
        p += step;
    }
    while (p < 1.0) {
        for (float q = 0; (double)q < 1.0 - (double)p; q += step) {
            const int x = (int)((double)(p * (float)x1 + q * (float)x2) + (1.0 - (double)p - (double)q) * x3);
            const int y = (int)((double)(p * (float)y1 + q * (float)y2) + (1.0 - (double)p - (double)q) * y3);
            if (x >= 0 && x < width && y >= 0 && y < height) {
                atomicMax(&mask[(size_t)y * width + x], val);
            }
        }
        /* the sums only grow: past 1 after 64 more steps iff past 1 at the
         * end of them */
        for (int k = 0; k < 64; ++k) {
            p += step;
        }
    }
}

/* the plane of every triangle (main.cpp:117-121) */
__global__ __launch_bounds__(256) void hpm_k_tri_planes(const int32_t* pts, const int32_t* tris, int ntris,
                                                        const float* depth, int width, hpm_camera ref, float factor,
                                                        float* planes)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= ntris) {
        return;
    }
    const int32_t p[6] = {pts[2 * tris[3 * t]], pts[2 * tris[3 * t] + 1], pts[2 * tris[3 * t + 1]],
                          pts[2 * tris[3 * t + 1] + 1], pts[2 * tris[3 * t + 2]], pts[2 * tris[3 * t + 2] + 1]};
    reinterpret_cast<float4*>(planes)[t] = plane_from_points(ref, p, depth, width, factor);
}

/* DepthsPredictSupplement (HPM.cpp:923-1097) as hpm_knn_supplement, one lane
 * per pixel no triangle covers: the k nearest support points in (squared
 * distance, index) order from a uniform grid (cells of `cell` px, point
 * indices ascending per cell; rings of cells until the next ring lies
 * strictly farther than the k-th), the first non-collinear triple of
 * `order`, its plane, kept when its depth is in range. */
__global__ __launch_bounds__(256) void hpm_k_knn(const int32_t* pts, int npoints, const int32_t* cell_start,
                                                 const int32_t* cell_pts, int cell, int gw, int gh,
                                                 const int32_t* order, int norder, int k, const uint32_t* raster,
                                                 const float* depth, int width, int height, hpm_camera ref,
                                                 float factor, float depth_min, float depth_max, uint8_t* knn_flag,
                                                 float* knn_planes)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (j >= width || i >= height) {
        return;
    }
    const size_t center = (size_t)i * width + j;
    knn_flag[center] = 0;
    if (raster[center] > 0u || npoints < 3) {
        return;
    }
    int64_t bd[32];
    int bi[32];
    int found = 0;
    const int kk = min(k, npoints);
    const int cx = min(j / cell, gw - 1), cy = min(i / cell, gh - 1);
    for (int r = 0;; ++r) {
        const int x_lo = cx - r, x_hi = cx + r, y_lo = cy - r, y_hi = cy + r;
        for (int gy = max(y_lo, 0); gy <= min(y_hi, gh - 1); ++gy) {
            for (int gx = max(x_lo, 0); gx <= min(x_hi, gw - 1); ++gx) {
                if (gx != x_lo && gx != x_hi && gy != y_lo && gy != y_hi) {
                    continue; /* inner cells: earlier rings */
                }
                const int c = gy * gw + gx;
                for (int e = cell_start[c]; e < cell_start[c + 1]; ++e) {
                    const int s = cell_pts[e];
                    const int64_t dx = pts[2 * s] - j, dy = pts[2 * s + 1] - i;
                    const int64_t d2 = dx * dx + dy * dy;
                    if (found == kk && (d2 > bd[kk - 1] || (d2 == bd[kk - 1] && s > bi[kk - 1]))) {
                        continue;
                    }
                    int pos = found < kk ? found : kk - 1;
                    while (pos > 0 && (bd[pos - 1] > d2 || (bd[pos - 1] == d2 && bi[pos - 1] > s))) {
                        bd[pos] = bd[pos - 1];
                        bi[pos] = bi[pos - 1];
                        --pos;
                    }
                    bd[pos] = d2;
                    bi[pos] = s;
                    if (found < kk) {
                        ++found;
                    }
                }
            }
        }
        if (x_lo <= 0 && y_lo <= 0 && x_hi >= gw - 1 && y_hi >= gh - 1) {
            break; /* every cell seen */
        }
        if (found == kk) {
            /* a point outside the rings so far lies beyond one of the sides
             * the rings have not reached the grid's edge on */
            int64_t gap = INT64_MAX;
            if (x_lo > 0) gap = min(gap, (int64_t)(j - x_lo * cell + 1));
            if (x_hi < gw - 1) gap = min(gap, (int64_t)((x_hi + 1) * cell - j));
            if (y_lo > 0) gap = min(gap, (int64_t)(i - y_lo * cell + 1));
            if (y_hi < gh - 1) gap = min(gap, (int64_t)((y_hi + 1) * cell - i));
            if (gap * gap > bd[kk - 1]) {
                break;
            }
        }
    }
    if (found < 3) {
        return;
    }
    for (int t = 0; t < norder; ++t) {
        if (order[3 * t + 2] >= found) {
            continue;
        }
        const int a = bi[order[3 * t]], b = bi[order[3 * t + 1]], c = bi[order[3 * t + 2]];
        const int64_t cr = (int64_t)(pts[2 * b] - pts[2 * a]) * (pts[2 * c + 1] - pts[2 * a + 1]) -
                           (int64_t)(pts[2 * b + 1] - pts[2 * a + 1]) * (pts[2 * c] - pts[2 * a]);
        if (cr != 0) {
            const int32_t tri[6] = {pts[2 * a], pts[2 * a + 1], pts[2 * b], pts[2 * b + 1], pts[2 * c], pts[2 * c + 1]};
            const float4 n4 = plane_from_points(ref, tri, depth, width, factor);
            const float d = depth_from_plane_factor(ref, n4, j, i, factor);
            if (d <= depth_max && d >= depth_min) {
                knn_flag[center] = 1;
                reinterpret_cast<float4*>(knn_planes)[center] = n4;
            }
            return;
        }
    }
}

/* The prior maps of one scale (main.cpp:124-162, 277-314 below full size,
 * 427-454 at full size): the triangle plane (mask idx + 1) or the KNN plane
 * (mask -233) where its depth passes the range gate. Below full size
 * (src_nd != null): world normal and depth per pixel for the JBU, the gate
 * relaxed to [0.6 dmin, 1.2 dmax] for triangles. At full size: camera-frame
 * planes and masks 1 (the KNN mask counts as active, Q12). */
__global__ __launch_bounds__(256) void hpm_k_prior_assemble(const uint32_t* raster, const float* tri_planes,
                                                            const uint8_t* knn_flag, const float* knn_planes,
                                                            int width, int height, hpm_camera ref, float factor,
                                                            float depth_min, float depth_max, float* src_nd,
                                                            float* planes, uint32_t* mask)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= width || y >= height) {
        return;
    }
    const size_t q = (size_t)y * width + x;
    const uint32_t r = raster[q];
    float4 plane = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool have = false, tri = false;
    if (r > 0u) {
        plane = reinterpret_cast<const float4*>(tri_planes)[r - 1u];
        have = tri = true;
    } else if (knn_flag[q]) {
        plane = reinterpret_cast<const float4*>(knn_planes)[q];
        have = true;
    }
    const float d = have ? depth_from_plane_factor(ref, plane, x, y, factor) : 0.0f;
    if (src_nd) {
        const bool ok = have && (tri ? (d <= depth_max * 1.2f && d >= depth_min * 0.6f)
                                     : (d <= depth_max && d >= depth_min));
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ok) {
            o.x = ref.R[0] * plane.x + ref.R[3] * plane.y + ref.R[6] * plane.z;
            o.y = ref.R[1] * plane.x + ref.R[4] * plane.y + ref.R[7] * plane.z;
            o.z = ref.R[2] * plane.x + ref.R[5] * plane.y + ref.R[8] * plane.z;
            o.w = d;
        }
        reinterpret_cast<float4*>(src_nd)[q] = o;
    } else {
        const bool ok = have && d <= depth_max && d >= depth_min;
        reinterpret_cast<float4*>(planes)[q] = ok ? plane : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        mask[q] = ok ? 1u : 0u;
    }
}

__global__ __launch_bounds__(256) void hpm_k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = v;
    }
}

/* ---- support points, their Delaunay triangulation and the KNN cells ---- */

/* one workgroup: out[0] = 0, out[i + 1] = in[0] + ... + in[i] (in may be
 * 0/1 flags of cand >= 0 when flags is set) */
__global__ __launch_bounds__(1024) void hpm_k_scan(const int32_t* in, int n, int flags, int32_t* out)
{
    __shared__ int32_t part[1024];
    const int t = (int)threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = min(n, t * per), e = min(n, b + per);
    int32_t sum = 0;
    for (int i = b; i < e; ++i) {
        sum += flags ? (in[i] >= 0 ? 1 : 0) : in[i];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = t > 0 ? part[t - 1] : 0;
    if (t == 0) {
        out[0] = 0;
    }
    for (int i = b; i < e; ++i) {
        run += flags ? (in[i] >= 0 ? 1 : 0) : in[i];
        out[i + 1] = run;
    }
}

/* the support points in block order (GetSupportPoints' scan): pts[k] for
 * the k-th non-empty block, and each block's point index (or -1) */
__global__ __launch_bounds__(256) void hpm_k_support_compact(const int32_t* cand, const int32_t* pos, int nb,
                                                             int32_t* pts, int32_t* bpoint)
{
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nb) {
        return;
    }
    const int32_t c = cand[b];
    if (c >= 0) {
        const int k = pos[b];
        pts[2 * k] = c & 0xffff;
        pts[2 * k + 1] = c >> 16;
        bpoint[b] = k;
    } else {
        bpoint[b] = -1;
    }
}

/* per pixel column x, the points with that x of least and greatest y
 * (their point indices, -1 if none): the convex hull's vertices are among
 * them */
__global__ __launch_bounds__(256) void hpm_k_hull_columns(const int32_t* pts, const int32_t* bpoint, int nby, int ws,
                                                          int32_t* col_lo, int32_t* col_hi)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (x >= ws) {
        return;
    }
    const int bc = x / 5;
    int lo = -1, hi = -1;
    for (int br = 0; br < nby; ++br) {
        const int i = bpoint[bc * nby + br];
        if (i >= 0 && pts[2 * i] == x) {
            if (lo < 0) {
                lo = i;
            }
            hi = i;
        }
    }
    col_lo[x] = lo;
    col_hi[x] = hi;
}

/* Andrew's monotone chain over the column extremes, which come sorted by
 * (x, y): the lower chain over the least-y points, the upper over the
 * greatest-y points (a lower-hull vertex is its column's lowest point, an
 * upper one its highest). The hull counter-clockwise (orient > 0) from the
 * least (x, y) point, as the host's chain over all points gives it. One
 * workgroup stages the extremes in LDS (points packed x | y << 16, x and y
 * below 2^15) and one thread runs the chain there. hull has room for
 * 2 ws + 2 points. */
constexpr int kHullMaxW = 4096;
__global__ __launch_bounds__(256) void hpm_k_hull_chain(const int32_t* pts, const int32_t* col_lo,
                                                        const int32_t* col_hi, int ws, int32_t* hull, int32_t* nhull)
{
    __shared__ int32_t s_lo[kHullMaxW], s_hi[kHullMaxW], s_st[kHullMaxW + 2];
    for (int x = (int)threadIdx.x; x < ws; x += 256) {
        const int lo = col_lo[x], hi = col_hi[x];
        s_lo[x] = lo < 0 ? -1 : (pts[2 * lo] | (pts[2 * lo + 1] << 16));
        s_hi[x] = hi < 0 ? -1 : (pts[2 * hi] | (pts[2 * hi + 1] << 16));
    }
    __syncthreads();
    if (threadIdx.x != 0) {
        return;
    }
    using hpm_dt::P2;
    auto un = [](int32_t v) { return P2{v & 0xffff, v >> 16}; };
    /* distinct points and the extreme ones */
    int n = 0, first = -1, last = -1;
    for (int x = 0; x < ws; ++x) {
        if (s_lo[x] >= 0) {
            n += s_hi[x] != s_lo[x] ? 2 : 1;
            if (first < 0) {
                first = x;
            }
            last = x;
        }
    }
    if (n < 3) {
        *nhull = 0;
        return;
    }
    int k = 0;
    auto push = [&](int32_t v, int floor_k) {
        while (k >= floor_k + 2 && hpm_dt::orient(un(s_st[k - 2]), un(s_st[k - 1]), un(v)) <= 0) {
            --k;
        }
        s_st[k++] = v;
    };
    /* lower chain: least (x, y) .. greatest (x, y) over the lowest points,
     * then the greatest column's highest point */
    for (int x = first; x <= last; ++x) {
        if (s_lo[x] >= 0) {
            push(s_lo[x], 0);
        }
    }
    if (s_hi[last] != s_lo[last]) {
        push(s_hi[last], 0);
    }
    const int lower = k;
    for (int e = 0; e < lower; ++e) {
        hull[2 * e] = s_st[e] & 0xffff;
        hull[2 * e + 1] = s_st[e] >> 16;
    }
    /* upper chain from the greatest point back to the least */
    k = 0;
    s_st[k++] = s_hi[last];
    for (int x = last - 1; x >= first; --x) {
        if (s_hi[x] >= 0) {
            push(s_hi[x], 0);
        }
    }
    if (s_lo[first] != s_hi[first]) {
        push(s_lo[first], 0);
    }
    /* the upper chain's ends are the lower chain's ends */
    int m = lower;
    for (int e = 1; e < k - 1; ++e) {
        hull[2 * m] = s_st[e] & 0xffff;
        hull[2 * m + 1] = s_st[e] >> 16;
        ++m;
    }
    *nhull = m;
}

/* the star walk's grid over the support blocks: 5-pixel cells, at most one
 * point each (blocks numbered column-major); nx / ny give the next occupied
 * cell along a row / column, so runs of empty cells cost one step */
struct BlockGrid {
    const int32_t* P;
    const int32_t* bpoint;
    const int32_t* H;
    const int32_t* nx; /* [y][x]: least x' >= x with a point in row y, else gw */
    const int32_t* ny; /* [x][y]: least y' >= y with a point in column x, else gh */
    int gw, gh, cs, nhull;

    __device__ hpm_dt::P2 pt(int i) const { return hpm_dt::P2{P[2 * i], P[2 * i + 1]}; }
    __device__ int cx(hpm_dt::P2 p) const { return p.x / 5; }
    __device__ int cy(hpm_dt::P2 p) const { return p.y / 5; }
    __device__ int cxd(double x) const { return (int)fmax(-1.0, fmin((double)gw, floor(x / 5.0))); }
    __device__ int cyd(double y) const { return (int)fmax(-1.0, fmin((double)gh, floor(y / 5.0))); }
    __device__ double rowy(int y) const { return 5.0 * y; }
    __device__ double colx(int x) const { return 5.0 * x; }
    __device__ hpm_dt::P2 hull(int k) const { return hpm_dt::P2{H[2 * k], H[2 * k + 1]}; }
    template <class F>
    __device__ void visit_row(int y, int xa, int xb, F&& f) const
    {
        for (int x = nx[y * gw + xa]; x <= xb; x = x + 1 < gw ? nx[y * gw + x + 1] : gw) {
            f(bpoint[x * gh + y]);
        }
    }
    template <class F>
    __device__ void visit_col(int x, int ya, int yb, F&& f) const
    {
        for (int y = ny[x * gh + ya]; y <= yb; y = y + 1 < gh ? ny[x * gh + y + 1] : gh) {
            f(bpoint[x * gh + y]);
        }
    }
};

/* the next-occupied-cell tables of BlockGrid: one thread per row (first
 * nby threads) and per column (next nbx threads) */
__global__ __launch_bounds__(256) void hpm_k_block_spans(const int32_t* bpoint, int nbx, int nby, int32_t* nx,
                                                         int32_t* ny)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < nby) {
        int next = nbx;
        for (int x = nbx - 1; x >= 0; --x) {
            if (bpoint[x * nby + t] >= 0) {
                next = x;
            }
            nx[t * nbx + x] = next;
        }
    } else if (t < nby + nbx) {
        const int x = t - nby;
        int next = nby;
        for (int y = nby - 1; y >= 0; --y) {
            if (bpoint[x * nby + y] >= 0) {
                next = y;
            }
            ny[x * nby + y] = next;
        }
    }
}

/* pass 0: the number of triangles each point emits (count[p + 1]), the
 * first kKeep of them kept in keep[p]; pass 1: the triangles, at the
 * offsets the scan of the counts gives, copied from keep, or walked again
 * for the rare points that emit more */
constexpr int kKeep = 8;
__global__ __launch_bounds__(256) void hpm_k_delaunay(const int32_t* pts, const int32_t* npts_p, const int32_t* bpoint,
                                                      const int32_t* nx, const int32_t* ny, int nbx, int nby,
                                                      const int32_t* hull, const int32_t* nhull, int pass,
                                                      int32_t* count, int32_t* keep, int32_t* tris, int max_tris,
                                                      int32_t* err)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int npts = *npts_p;
    if (p >= nbx * nby) {
        return;
    }
    if (p >= npts) { /* the scan runs over one count per block */
        if (pass == 0) {
            count[p + 1] = 0;
        }
        return;
    }
    BlockGrid g;
    g.P = pts;
    g.bpoint = bpoint;
    g.nx = nx;
    g.ny = ny;
    g.H = hull;
    g.gw = nbx;
    g.gh = nby;
    g.cs = 5;
    g.nhull = *nhull;
    if (npts < 3) {
        if (pass == 0) {
            count[p + 1] = 0;
        }
        return;
    }
    int32_t* kp = keep + (size_t)p * 2 * kKeep;
    if (pass == 0) {
        int c = 0;
        const int r = hpm_dt::star(g, p, 4 * npts + 16, [&](int a, int b) {
            if (c < kKeep) {
                kp[2 * c] = a;
                kp[2 * c + 1] = b;
            }
            ++c;
        });
        count[p + 1] = c;
        if (r < 0) {
            atomicOr(err, 1);
        }
    } else {
        int at = count[p];
        const int c = count[p + 1] - at;
        auto put = [&](int a, int b) {
            if (at < max_tris) {
                tris[3 * at] = p;
                tris[3 * at + 1] = a;
                tris[3 * at + 2] = b;
            } else {
                atomicOr(err, 2);
            }
            ++at;
        };
        if (c <= kKeep) {
            for (int k = 0; k < c; ++k) {
                put(kp[2 * k], kp[2 * k + 1]);
            }
        } else {
            hpm_dt::star(g, p, 4 * npts + 16, put);
        }
    }
}

/* the rasteriser's running index: every support point lies in the image,
 * so every triangle is counted (imageRC.contains, main.cpp:97) */
__global__ __launch_bounds__(256) void hpm_k_iota(int32_t* p, const int32_t* n_p)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < *n_p) {
        p[i] = i;
    }
}

/* The KNN search grid: cells of `cell` px, point indices ascending within a
 * cell. A cell's points come from the blocks it meets, visited column-major
 * (ascending point index); pass 0 counts, pass 1 writes at the scanned
 * offsets. */
__global__ __launch_bounds__(256) void hpm_k_knn_cells(const int32_t* pts, const int32_t* bpoint, int nbx, int nby,
                                                       int cell, int gw, int gh, int pass, int32_t* count,
                                                       int32_t* items)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= gw * gh) {
        return;
    }
    const int cx = c % gw, cy = c / gw;
    const int bc0 = cx * cell / 5, bc1 = min(nbx - 1, (cx * cell + cell - 1) / 5);
    const int br0 = cy * cell / 5, br1 = min(nby - 1, (cy * cell + cell - 1) / 5);
    int n = 0;
    int at = pass == 1 ? count[c] : 0;
    for (int bc = bc0; bc <= bc1; ++bc) {
        for (int br = br0; br <= br1; ++br) {
            const int i = bpoint[bc * nby + br];
            if (i < 0) {
                continue;
            }
            const int x = pts[2 * i], y = pts[2 * i + 1];
            if (min(x / cell, gw - 1) != cx || min(y / cell, gh - 1) != cy) {
                continue;
            }
            if (pass == 1) {
                items[at++] = i;
            }
            ++n;
        }
    }
    if (pass == 0) {
        count[c + 1] = n;
    }
}

extern "C" int hpm_launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, void* stream)
{
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(hpm_k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, p, n,
                       value);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_prior_maps(const float* hyp4, const float* cost, int W, int H, float* depth_s,
                                     float* cost_s, int ws, int hs, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_prior_maps, dim3((unsigned)((ws + 63) / 64), (unsigned)((hs + 3) / 4), 1), blk, 0,
                       (hipStream_t)stream, hyp4, cost, W, H, depth_s, cost_s, ws, hs);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_support_blocks(const float* cost, int w, int h, int32_t* cand, void* stream)
{
    const int nb = ((w + 4) / 5) * ((h + 4) / 5);
    hipLaunchKernelGGL(hpm_k_support_blocks, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       cost, w, h, cand);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_raster(const int32_t* pts, const int32_t* tris, const int32_t* tri_idx, int ntris,
                                 uint32_t* mask, int width, int height, void* stream)
{
    const int64_t n = (int64_t)width * height;
    hipLaunchKernelGGL(hpm_k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, mask, n,
                       0u);
    if (ntris > 0) {
        hipLaunchKernelGGL(hpm_k_raster, dim3((unsigned)((ntris + 3) / 4)), dim3(256), 0, (hipStream_t)stream, pts,
                           tris, tri_idx, ntris, mask, width, height);
    }
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_tri_planes(const int32_t* pts, const int32_t* tris, int ntris, const float* depth,
                                     int width, const hpm_camera* ref, float factor, float* planes, void* stream)
{
    if (ntris > 0) {
        hipLaunchKernelGGL(hpm_k_tri_planes, dim3((unsigned)((ntris + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, pts, tris, ntris, depth, width, *ref, factor, planes);
    }
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_knn(const int32_t* pts, int npoints, const int32_t* cell_start, const int32_t* cell_pts,
                              int cell, int gw, int gh, const int32_t* order, int norder, int k,
                              const uint32_t* raster, const float* depth, int width, int height,
                              const hpm_camera* ref, float factor, float depth_min, float depth_max,
                              uint8_t* knn_flag, float* knn_planes, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_knn, dim3((unsigned)((width + 63) / 64), (unsigned)((height + 3) / 4), 1), blk, 0,
                       (hipStream_t)stream, pts, npoints, cell_start, cell_pts, cell, gw, gh, order, norder, k, raster,
                       depth, width, height, *ref, factor, depth_min, depth_max, knn_flag, knn_planes);
    return (int)hipGetLastError();
}

extern "C" int hpm_launch_prior_assemble(const uint32_t* raster, const float* tri_planes, const uint8_t* knn_flag,
                                         const float* knn_planes, int width, int height, const hpm_camera* ref,
                                         float factor, float depth_min, float depth_max, float* src_nd,
                                         float* planes, uint32_t* mask, void* stream)
{
    const dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(hpm_k_prior_assemble, dim3((unsigned)((width + 63) / 64), (unsigned)((height + 3) / 4), 1),
                       blk, 0, (hipStream_t)stream, raster, tri_planes, knn_flag, knn_planes, width, height, *ref,
                       factor, depth_min, depth_max, src_nd, planes, mask);
    return (int)hipGetLastError();
}

/* The support points of the block candidates, their Delaunay triangulation
 * and the KNN cells, all on the device. ws_i32 has room for the scans, the
 * span tables, the kept triangles and the hull (21 nb + 10 ws + 16 int32).
 * nums receives (npts, ntris, error). Widths up to 4096. */
extern "C" int hpm_launch_support_delaunay(const int32_t* cand, int ws, int hs, int cell, int gw, int gh,
                                           int32_t* pts, int32_t* tris, int32_t* tri_idx, int32_t* cell_start,
                                           int32_t* cell_pts, int32_t* ws_i32, int32_t* nums, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int nbx = (ws + 4) / 5, nby = (hs + 4) / 5, nb = nbx * nby;
    int32_t* pos = ws_i32;                /* nb + 1 */
    int32_t* bpoint = pos + nb + 1;       /* nb */
    int32_t* col_lo = bpoint + nb;        /* ws */
    int32_t* col_hi = col_lo + ws;        /* ws */
    int32_t* hull = col_hi + ws;          /* 2 (4 ws + 2): the chain's stack over <= 2 ws points */
    int32_t* nhull = hull + 2 * (4 * ws + 2);
    int32_t* tcount = nhull + 1;          /* nb + 1 (points <= blocks) */
    int32_t* nxt = tcount + nb + 1;       /* nb */
    int32_t* nyt = nxt + nb;              /* nb */
    int32_t* keep = nyt + nb;             /* 2 kKeep nb */
    int32_t* err = nums + 2;
    hipLaunchKernelGGL(hpm_k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)err, (int64_t)1, 0u);
    hipLaunchKernelGGL(hpm_k_scan, dim3(1), dim3(1024), 0, s, cand, nb, 1, pos);
    hipLaunchKernelGGL(hpm_k_support_compact, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, cand, pos, nb, pts,
                       bpoint);
    hipLaunchKernelGGL(hpm_k_hull_columns, dim3((unsigned)((ws + 255) / 256)), dim3(256), 0, s, pts, bpoint, nby, ws,
                       col_lo, col_hi);
    if (ws > kHullMaxW) {
        return (int)hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(hpm_k_hull_chain, dim3(1), dim3(256), 0, s, pts, col_lo, col_hi, ws, hull, nhull);
    hipLaunchKernelGGL(hpm_k_block_spans, dim3((unsigned)((nbx + nby + 255) / 256)), dim3(256), 0, s, bpoint, nbx,
                       nby, nxt, nyt);
    const int32_t* npts_p = pos + nb;
    /* at most nb points: the grids cover every block, threads past npts return */
    const unsigned gp = (unsigned)((nb + 255) / 256);
    /* a triangulation of n points has at most 2 n - 5 triangles */
    const int max_tris = 2 * nb;
    hipLaunchKernelGGL(hpm_k_delaunay, dim3(gp), dim3(256), 0, s, pts, npts_p, bpoint, nxt, nyt, nbx, nby, hull,
                       nhull, 0, tcount, keep, tris, max_tris, err);
    hipLaunchKernelGGL(hpm_k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)tcount, (int64_t)1, 0u);
    hipLaunchKernelGGL(hpm_k_scan, dim3(1), dim3(1024), 0, s, tcount + 1, nb, 0, tcount);
    hipLaunchKernelGGL(hpm_k_delaunay, dim3(gp), dim3(256), 0, s, pts, npts_p, bpoint, nxt, nyt, nbx, nby, hull,
                       nhull, 1, tcount, keep, tris, max_tris, err);
    hipLaunchKernelGGL(hpm_k_iota, dim3((unsigned)((max_tris + 255) / 256)), dim3(256), 0, s, tri_idx, tcount + nb);
    hipLaunchKernelGGL(hpm_k_knn_cells, dim3((unsigned)((gw * gh + 255) / 256)), dim3(256), 0, s, pts, bpoint, nbx,
                       nby, cell, gw, gh, 0, cell_start, cell_pts);
    hipLaunchKernelGGL(hpm_k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)cell_start, (int64_t)1, 0u);
    hipLaunchKernelGGL(hpm_k_scan, dim3(1), dim3(1024), 0, s, cell_start + 1, gw * gh, 0, cell_start);
    hipLaunchKernelGGL(hpm_k_knn_cells, dim3((unsigned)((gw * gh + 255) / 256)), dim3(256), 0, s, pts, bpoint, nbx,
                       nby, cell, gw, gh, 1, cell_start, cell_pts);
    hipError_t e = hipMemcpyAsync(nums, npts_p, 4, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(nums + 1, tcount + nb, 4, hipMemcpyDeviceToDevice, s);
    }
    return e != hipSuccess ? (int)e : (int)hipGetLastError();
}
