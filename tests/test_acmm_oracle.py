"""Pins the CPU oracle's ACMM+NESP variant (hpm_params.acmm / hierarchy,
ACMM+NESP/windows/ACMM.cu) on the CPU.

  - without geometric consistency the variant differs from HPM-MVS only in
    what never triggers there (the NaN fallback): a 3-iteration run is
    bit-identical to the HPM-MVS run;
  - RunPatchMatch runs max_iterations iterations (ACMM.cu:1836): a run with 2
    equals init + 2 staged iterations + finalize + filter, not 3;
  - the geometric cost against a constant source depth map in closed form
    (pure x translation: reprojection error |fx tx (1/D - 1/D')|), capped at
    3 in ACMM and 5 in HPM-MVS, the cap for a zero source depth;
  - the JBU (window, clamps, spatial and range weights, normalisation)
    against a pure-Python loop restatement on a small map;
  - the hierarchy initialisation: depth = the JBU depth, normal = R times
    the JBU normal; pre-cost = the top-k cost of the plane (0, 0, 0, depth)
    against the float64 cost vectors;
  - the hierarchy commit rule (ACMM.cu:1619-1628): after a half-iteration,
    every pixel whose cost did not beat its pre-cost by 0.1 kept its
    hypothesis.
"""
import ctypes as C

import numpy as np
import pytest


def _acmm_params(cpu, cams, **kw):
    p = cpu.relaxed_params(cams)
    p.acmm = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_acmm_plain_equals_hpm(cpu, c1_small):
    cfg, (cams, images, gd, gn, low, order) = c1_small
    a = cpu.Oracle(cams, images, cpu.relaxed_params(cams), salt=3)
    a.run_patchmatch()
    b = cpu.Oracle(cams, images, _acmm_params(cpu, cams, max_iterations=3), salt=3)
    b.run_patchmatch()
    assert np.array_equal(a.hyp.view(np.uint32), b.hyp.view(np.uint32))
    assert np.array_equal(a.cost.view(np.uint32), b.cost.view(np.uint32))
    assert np.array_equal(a.sel, b.sel)
    assert a.n_eval == b.n_eval


def test_acmm_runs_max_iterations(cpu, c1_small):
    cfg, (cams, images, gd, gn, low, order) = c1_small
    run = cpu.Oracle(cams, images, _acmm_params(cpu, cams, max_iterations=2), salt=1)
    run.run_patchmatch()
    staged = {}
    for iters in (2, 3):
        o = cpu.Oracle(cams, images, _acmm_params(cpu, cams, max_iterations=2), salt=1)
        o.init()
        for it in range(iters):
            o.half_iteration(it, 0)
            o.half_iteration(it, 1)
        o.finalize()
        o.filter(0)
        o.filter(1)
        staged[iters] = o.hyp.copy()
    assert np.array_equal(run.hyp.view(np.uint32), staged[2].view(np.uint32))
    assert not np.array_equal(run.hyp.view(np.uint32), staged[3].view(np.uint32))


def _cam(hpm, fx, cx, cy, W, H, tx=0.0, dmin=1.0, dmax=10.0):
    K = [fx, 0, cx, 0, fx, cy, 0, 0, 1]
    return hpm.camera_from_arrays(K, np.eye(3), [tx, 0.0, 0.0], W, H, dmin, dmax)


@pytest.mark.parametrize("acmm,cap", [(0, 5.0), (1, 3.0)])
def test_geom_cost_closed_form_and_cap(hpm, cpu, acmm, cap):
    W, H, fx, tx, D = 64, 48, 60.0, 0.1, 4.0
    cams = [_cam(hpm, fx, 32.0, 24.0, W, H), _cam(hpm, fx, 32.0, 24.0, W, H, tx=tx)]
    images = [np.full((H, W), 100.0, np.float32)] * 2
    L = cpu.lib()
    L.hpm_oracle_geom_cost.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int,
                                       C.POINTER(C.c_float)]
    ys, xs = np.mgrid[8:40:5, 8:40:5]
    pix = np.ascontiguousarray(np.stack([xs.ravel(), ys.ravel()], 1), np.int32)
    planes = np.tile(np.array([0.0, 0.0, -1.0, D], np.float32), (len(pix), 1))
    for d_src in (4.0, 5.0, 2.0, 1.0, 0.5, 0.0):
        depths = [np.full((H, W), D, np.float32), np.full((H, W), d_src, np.float32)]
        p = cpu.relaxed_params(cams)
        p.acmm = acmm
        p.geom_consistency = 1
        o = cpu.Oracle(cams, images, p, depths=depths)
        out = np.empty(len(pix), np.float32)
        rc = L.hpm_oracle_geom_cost(C.byref(o.pb), 1, planes.ctypes.data_as(C.POINTER(C.c_float)),
                                    pix.ctypes.data_as(C.POINTER(C.c_int32)), len(pix),
                                    out.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0
        want = cap if d_src == 0.0 else min(cap, abs(fx * tx * (1.0 / D - 1.0 / d_src)))
        assert np.allclose(out, want, atol=2e-4), (d_src, out[:4], want)


def _jbu_python(guide, src_nd, scale_i):
    """JBU_cu (HPM.cu:1984-2052 / ACMM.cu:1862-1919) as plain loops, float64"""
    h, w = guide.shape
    sh, sw = src_nd.shape[:2]
    scale = sw / w
    nn = (scale_i * scale_i + 1) // 2
    depth = np.zeros((h, w))
    normal = np.zeros((h, w, 3))
    for py in range(h):
        for px in range(w):
            oy, ox = np.float32(py * np.float32(scale)), np.float32(px * np.float32(scale))
            ref = guide[py, px]
            nf = dt = 0.0
            nt = np.zeros(3)
            for j in range(-nn, nn + 1):
                ry = min(max(int(oy + j), 0), sh - 1)
                rys = min(max(py + j, 0), h - 1)
                for i in range(-nn, nn + 1):
                    rx = min(max(int(ox + i), 0), sw - 1)
                    nb = guide[rys, min(max(px + i, 0), w - 1)]
                    sg = np.exp(-((ox - rx) ** 2 + (oy - ry) ** 2) / (2 * 0.5 * 0.5))
                    rg = np.exp(-((ref - nb) ** 2) / (2 * 25.5 * 25.5))
                    t = sg * rg
                    nf += t
                    dt += src_nd[ry, rx, 3] * t
                    nt += src_nd[ry, rx, :3] * t
            depth[py, px] = dt / nf
            n = nt / nf
            normal[py, px] = n / np.linalg.norm(n)
    return depth, normal


def test_jbu_against_python_loops(cpu):
    rng = np.random.default_rng(5)
    h, w, sh, sw = 14, 18, 7, 9
    guide = rng.uniform(0, 255, (h, w)).astype(np.float32)
    src = np.zeros((sh, sw, 4), np.float32)
    n = rng.normal(size=(sh, sw, 3))
    n[..., 2] = -np.abs(n[..., 2]) - 1.0
    src[..., :3] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    src[..., 3] = rng.uniform(2, 6, (sh, sw))
    d, nn = cpu.jbu(guide, src)
    wd, wn = _jbu_python(guide.astype(np.float64), src.astype(np.float64), 2)
    assert np.allclose(d, wd, rtol=2e-6, atol=0)
    assert np.allclose(nn[..., :3], wn, atol=2e-6)


@pytest.fixture(scope="module")
def hierarchy_case(hpm, cpu):
    """A plain ACMM run at 80x60 and the hierarchy initialisation at
    160x120 from it."""
    cfg = hpm.scene_config(seed=12, width=160, height=120, fx=125.0, fy=125.0, cx=80.0, cy=60.0)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    order = [2, 1, 3, 0, 4]
    from oracle import pipeline

    fine = [cams[i] for i in order]
    fimg = [imgs[i].astype(np.float32) for i in order]
    coarse = [pipeline._scale_view(c, im, 80) for c, im in zip(fine, fimg)]
    cc = [c for c, _ in coarse]
    ci = [im for _, im in coarse]
    o = cpu.Oracle(cc, ci, _acmm_params(cpu, cc, max_iterations=3), salt=2)
    o.run_patchmatch()
    h = cpu.Oracle(fine, fimg, _acmm_params(cpu, fine, max_iterations=3), salt=2)
    h.set_hierarchy(o.hyp)
    h.pb.call_index = 16
    h.init()
    return fine, fimg, o.hyp.copy(), h


def test_hierarchy_init(cpu, hierarchy_case):
    fine, fimg, coarse, h = hierarchy_case
    H, W = h.cost.shape
    d, n = cpu.jbu(fimg[0], coarse)
    R = np.array(fine[0].R, np.float32).reshape(3, 3)
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    want = np.stack([(R[0, 0] * x + R[0, 1] * y) + R[0, 2] * z, (R[1, 0] * x + R[1, 1] * y) + R[1, 2] * z,
                     (R[2, 0] * x + R[2, 1] * y) + R[2, 2] * z], -1)
    assert np.array_equal(h.hyp[..., :3].view(np.uint32), want.astype(np.float32).view(np.uint32))
    # the plane's depth at each pixel is the JBU depth
    K = fine[0].K
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    pl = h.hyp.astype(np.float64)
    depth = -pl[..., 3] * K[0] / ((xs - K[2]) * pl[..., 0] + (K[0] / K[4]) * (ys - K[5]) * pl[..., 1] +
                                  K[0] * pl[..., 2])
    assert np.allclose(depth, d, rtol=1e-4)
    # pre-cost: the top-k mean cost of the plane (0, 0, 0, depth), float64 anchor
    rng = np.random.default_rng(1)
    idx = rng.integers(0, W * H, 300)
    pix = np.stack([idx % W, idx // W], 1)
    planes = np.zeros((len(idx), 4), np.float32)
    planes[:, 3] = d.reshape(-1)[idx]
    cv = h.cost_vectors(planes, pix, fp64=True)
    top = []
    for c in cv:
        valid = int((c < 2.0).sum())
        k = min(valid, 4)
        top.append(2.0 if k == 0 else float(np.sort(c)[:k].mean()))
    assert np.allclose(h.pre_cost.reshape(-1)[idx], top, atol=1e-4)
    assert 0.05 < np.mean(h.pre_cost < 2.0)


def test_hierarchy_commit_rule(cpu, hierarchy_case):
    fine, fimg, coarse, h = hierarchy_case
    hyp0, cost0, pre = h.hyp.copy(), h.cost.copy(), h.pre_cost.copy()
    h.half_iteration(0, 0)
    H, W = h.cost.shape
    ys, xs = np.mgrid[0:H, 0:W]
    black = (xs + ys) % 2 == 0
    kept = black & ~(h.cost < pre - np.float32(0.1))
    committed = black & (h.cost < pre - np.float32(0.1))
    assert kept.sum() > 0 and committed.sum() > 0
    assert np.array_equal(h.hyp[kept].view(np.uint32), hyp0[kept].view(np.uint32))
    assert np.array_equal(h.hyp[~black].view(np.uint32), hyp0[~black].view(np.uint32))
    changed = black & np.any(h.hyp != hyp0, axis=-1)
    assert changed.sum() > 0 and not np.any(changed & kept)
