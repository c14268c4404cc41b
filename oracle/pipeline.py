"""CPU restatement of the HPM-MVS pass schedule (main.cpp:35-499, 636-674)
and of the ACMM+NESP coarse-to-fine schedule (ACMM+NESP/windows/main.cpp:
35-378).

TEST INFRASTRUCTURE: the checker of hpm_pipeline_run. It runs every
RunPatchMatch and JBU on the CPU oracle and builds the planar prior with the
library's host functions (support points, Delaunay, rasterisation, plane
fits, KNN supplement; pinned by tests/test_prior.py), in the same order and
with the same random streams as the device schedule.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import cpu

_hpm = cpu._hpm


def _jbu_prior(guide, src_nd, ref_cam, dmin, dmax, threads):
    L = cpu.lib()
    L.hpm_oracle_jbu_prior.argtypes = [cpu._FP, C.c_int, C.c_int, cpu._FP, C.c_int, C.c_int,
                                       C.POINTER(_hpm.HpmCamera), C.c_float, C.c_float, cpu._FP, cpu._UP, C.c_int]
    L.hpm_oracle_jbu_prior.restype = C.c_int
    g = np.ascontiguousarray(guide, np.float32)
    s = np.ascontiguousarray(src_nd, np.float32)
    h, w = g.shape
    sh, sw = s.shape[:2]
    planes = np.zeros((h, w, 4), np.float32)
    mask = np.zeros((h, w), np.uint32)
    rc = L.hpm_oracle_jbu_prior(cpu._f(g), w, h, cpu._f(s), sw, sh, C.byref(ref_cam), dmin, dmax, cpu._f(planes),
                                mask.ctypes.data_as(cpu._UP), threads)
    return None if rc == 1 else (planes, mask)


def _depth_from_plane_factor(K, planes, xs, ys, f):
    """GetDepthFromPlaneParam_factor (HPM.cpp:846-849) in float32, the
    expression order of hpm_depth_from_plane_factor."""
    f32 = np.float32
    K0, K2, K4, K5 = f32(K[0]), f32(K[2]), f32(K[4]), f32(K[5])
    f = f32(f)
    p0, p1, p2, p3 = planes[..., 0], planes[..., 1], planes[..., 2], planes[..., 3]
    num = (-p3) * (K0 * f)
    t0 = (xs.astype(f32) - K2 * f) * p0
    t1 = ((K0 / K4) * (ys.astype(f32) - K5 * f)) * p1
    t2 = (K0 * f) * p2
    with np.errstate(all="ignore"):
        return num / ((t0 + t1) + t2)


def _to_world(R, n):
    f32 = np.float32
    R = [f32(v) for v in R]
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    return np.stack([(R[0] * x + R[3] * y) + R[6] * z, (R[1] * x + R[4] * y) + R[7] * z,
                     (R[2] * x + R[5] * y) + R[8] * z], -1)


def _prior_scale(ref_cam, depth, cost, f, k, dmin, dmax, threads):
    h, w = depth.shape
    pts = _hpm.support_points(cost)
    tris = _hpm.delaunay(pts) if len(pts) >= 3 else np.zeros((0, 3), np.int32)
    mask = _hpm.rasterize_triangles(pts, tris, w, h)
    tri_planes = np.array([_hpm.plane_from_points(ref_cam, pts[t], depth, f) for t in tris],
                          np.float32).reshape(-1, 4)
    mask, knn = _hpm.knn_supplement(ref_cam, pts, depth, f, k, dmin, dmax, mask, threads)
    return mask, tri_planes, knn


def prior_scale_maps(ref_cam, depth, cost, guide, f, k, dmin, dmax, threads=0):
    """The planar prior of one scale from a finished pass's depth and cost
    (main.cpp:84-162 at f = 0.25, 238-314 at 0.5, 387-454 at 1), built with
    the library's host functions: (camera-frame planes (H, W, 4), masks
    (H, W)) at full size."""
    f32 = np.float32
    H, W = depth.shape
    if f < 1.0:
        ws, hs = int(np.floor(W * f + 0.5)), int(np.floor(H * f + 0.5))
        ds = _hpm.resize_bilinear(depth, ws, hs)
        cs_ = _hpm.resize_bilinear(cost, ws, hs)
        mask, tri_planes, knn = _prior_scale(ref_cam, ds, cs_, f, k, dmin, dmax, threads)
        ys, xs = np.mgrid[0:hs, 0:ws]
        src_nd = np.zeros((hs, ws, 4), np.float32)
        tri = mask > 0
        pl = np.zeros((hs, ws, 4), np.float32)
        if len(tri_planes):
            pl[tri] = tri_planes[mask[tri].astype(np.int64) - 1]
        d_tri = _depth_from_plane_factor(ref_cam.K, pl, xs, ys, f)
        ok_tri = tri & (d_tri <= dmax * f32(1.2)) & (d_tri >= dmin * f32(0.6))
        knn_m = mask < -200
        d_knn = _depth_from_plane_factor(ref_cam.K, knn, xs, ys, f)
        ok_knn = knn_m & (d_knn <= dmax) & (d_knn >= dmin)
        src_nd[ok_tri, 3] = d_tri[ok_tri]
        src_nd[ok_knn, 3] = d_knn[ok_knn]
        src_nd[ok_tri, :3] = _to_world(ref_cam.R, pl[ok_tri])
        src_nd[ok_knn, :3] = _to_world(ref_cam.R, knn[ok_knn])
        res = _jbu_prior(guide, src_nd, ref_cam, dmin, dmax, threads)
        if res is None:
            return np.zeros((H, W, 4), np.float32), np.zeros((H, W), np.uint32)
        return res
    mask, tri_planes, knn = _prior_scale(ref_cam, depth, cost, 1.0, k, dmin, dmax, threads)
    ys, xs = np.mgrid[0:H, 0:W]
    pl = np.zeros((H, W, 4), np.float32)
    tri = mask > 0
    if len(tri_planes):
        pl[tri] = tri_planes[mask[tri].astype(np.int64) - 1]
    knn_m = mask < -200
    pl[knn_m] = knn[knn_m]
    d = _depth_from_plane_factor(ref_cam.K, pl, xs, ys, 1.0)
    act = (tri | knn_m) & (d <= dmax) & (d >= dmin)
    planes = np.where(act[..., None], pl, f32(0)).astype(np.float32)
    return planes, act.astype(np.uint32)


def run_cpu_pipeline(cams, images, problems, opt, threads=0):
    """cams/images indexed by view; problems: [(ref, [sources])]; opt: an
    HpmPipelineOptions. Returns {view: (hyp_world (H, W, 4), cost (H, W))}
    plus the plain and geometric depth maps."""
    f32 = np.float32
    params = opt.params
    state, d_plain, d_geom = {}, {}, {}
    for ref, srcs in problems:
        order = [ref] + list(srcs)
        cs = [cams[v] for v in order]
        ims = [images[v] for v in order]
        p = cpu.relaxed_params(cs, params)
        p.planar_prior = p.geom_consistency = p.multi_geometry = 0
        o = cpu.Oracle(cs, ims, p, salt=ref, threads=threads)
        o.run_patchmatch()
        W, H = cs[0].width, cs[0].height
        dmin, dmax = f32(p.depth_min), f32(p.depth_max)
        if opt.planar_prior:
            o.pb.params.planar_prior = 1
            k, m = 0, max(W, H)
            while m > opt.scale_bound:
                m //= 2
                k += 1
            for i in range(2 - k, 3):
                f = (0.25, 0.5, 1.0)[i]
                planes, pmask = prior_scale_maps(cs[0], np.ascontiguousarray(o.hyp[..., 3]), o.cost.copy(), ims[0], f,
                                                 params.knn_k, dmin, dmax, threads)
                o.set_prior(planes, pmask)
                o.run_patchmatch()
        state[ref] = (o.hyp.copy(), o.cost.copy())
        d_plain[ref] = np.ascontiguousarray(o.hyp[..., 3]).copy()
    if opt.geom_consistency:
        for g in range(params.geom_iterations):
            for ref, srcs in problems:
                order = [ref] + list(srcs)
                cs = [cams[v] for v in order]
                ims = [images[v] for v in order]
                maps = d_geom if g > 0 else d_plain
                deps = [maps[v] for v in order]
                p = cpu.relaxed_params(cs, params)
                p.planar_prior = 0
                p.geom_consistency = 1
                p.multi_geometry = 1 if g > 0 else 0
                p.max_iterations = 2
                o = cpu.Oracle(cs, ims, p, depths=deps, salt=ref, threads=threads)
                o.hyp[...] = state[ref][0]
                o.cost[...] = state[ref][1]
                o.pb.call_index = 8 + g
                o.run_patchmatch()
                state[ref] = (o.hyp.copy(), o.cost.copy())
                d_geom[ref] = np.ascontiguousarray(o.hyp[..., 3]).copy()
    return state, d_plain, d_geom


def _scale_view(cam, image, cur):
    """ACMM::InuputInitialization's rescale (ACMM.cpp:497-523) in float32,
    with the library's bilinear resize (the device schedule's)."""
    f32 = np.float32
    W, H = cam.width, cam.height
    c = _hpm.HpmCamera()
    C.pointer(c)[0] = cam
    if W <= cur and H <= cur:
        return c, np.ascontiguousarray(image, np.float32)
    f = min(f32(cur) / f32(W), f32(cur) / f32(H))
    nw = int(np.floor(np.float64(f32(W) * f) + 0.5))
    nh = int(np.floor(np.float64(f32(H) * f) + 0.5))
    out = _hpm.resize_bilinear(np.ascontiguousarray(image, np.float32), nw, nh)
    sx, sy = f32(nw) / f32(W), f32(nh) / f32(H)
    c.K[0] = float(f32(c.K[0]) * sx)
    c.K[2] = float(f32(c.K[2]) * sx)
    c.K[4] = float(f32(c.K[4]) * sy)
    c.K[5] = float(f32(c.K[5]) * sy)
    c.width, c.height = nw, nh
    return c, out


def run_cpu_acmm(cams, images, problems, opt, threads=0):
    """The ACMM+NESP levels (main.cpp:302-374) on the CPU oracle, in the
    device schedule's order and random streams (hpm_acmm_run). opt: an
    HpmAcmmOptions. Returns (state {view: (hyp_world, cost)}, first-pass
    depths, geometric depths, level cameras) of the last level."""
    params = opt.params
    prob_of = {r: i for i, (r, _) in enumerate(problems)}
    max_size, left, cur = [], [], []
    for r, _ in problems:
        m = min(max(cams[r].width, cams[r].height), params.max_image_size)
        max_size.append(m)
        k = 0
        while m > opt.scale_bound:
            m //= 2
            k += 1
        left.append(k)
        cur.append(max_size[-1])
    prev = None
    for step in range(max(left) + 1):
        for i in range(len(problems)):
            if left[i] >= 0:
                cur[i] = int(max_size[i] / 2.0 ** left[i])
                left[i] -= 1
        lc, li = [], []
        for v in range(len(cams)):
            size = cur[prob_of[v]] if v in prob_of else max(cams[v].width, cams[v].height)
            c, im = _scale_view(cams[v], images[v], size)
            lc.append(c)
            li.append(im)
        state, d_first, d_geom = {}, {}, {}
        for ref, srcs in problems:
            order = [ref] + list(srcs)
            cs = [lc[v] for v in order]
            p = cpu.relaxed_params(cs, params)
            p.planar_prior = p.geom_consistency = p.multi_geometry = p.hierarchy = 0
            o = cpu.Oracle(cs, [li[v] for v in order], p, salt=ref, threads=threads)
            o.pb.call_index = 16 * step
            if prev is not None:
                o.set_hierarchy(prev[ref][0])
            o.run_patchmatch()
            state[ref] = (o.hyp.copy(), o.cost.copy())
            d_first[ref] = np.ascontiguousarray(o.hyp[..., 3]).copy()
        for g in range(opt.geom_iterations):
            for ref, srcs in problems:
                order = [ref] + list(srcs)
                cs = [lc[v] for v in order]
                maps = d_geom if g > 0 else d_first
                p = cpu.relaxed_params(cs, params)
                p.planar_prior = p.hierarchy = 0
                p.geom_consistency = 1
                p.multi_geometry = 1 if g > 0 else 0
                p.max_iterations = 2
                o = cpu.Oracle(cs, [li[v] for v in order], p, depths=[maps[v] for v in order], salt=ref,
                               threads=threads)
                o.hyp[...] = state[ref][0]
                o.cost[...] = state[ref][1]
                o.pb.call_index = 16 * step + 8 + g
                o.run_patchmatch()
                state[ref] = (o.hyp.copy(), o.cost.copy())
                d_geom[ref] = np.ascontiguousarray(o.hyp[..., 3]).copy()
        prev = state
    return state, d_first, d_geom, lc
