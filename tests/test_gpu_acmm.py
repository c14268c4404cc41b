"""The ACMM+NESP variant on the GPU (include/hpm_acmm.h, hpm_params.acmm /
hierarchy, hpm_set_hierarchy_input) against the CPU oracle.

- One hierarchy RunPatchMatch through the C-ABI (a 160x120 problem started
  from an 80x60 ACMM run): the device maps equal the oracle's bit for bit.
- A geometric ACMM pass (cap 3, weight 0.2, 2 iterations): bit for bit.
- The whole coarse-to-fine schedule (hpm_acmm_run) on a 4-view 320x240
  scene with the scale bound lowered to 200 (levels 160x120 and 320x240,
  two geometric passes each) against oracle/pipeline.py: run_cpu_acmm; bar
  as for the HPM-MVS schedule (at least 99% of pixels within 1% depth and
  5 degrees normal; expected identical).
- `hpm_mvs <folder> --acmm` (device schedule, and with --class-api the
  reference's main over the `hpm::ACMM` class; one level and two):
  writes ACMM_NESP/2333_%08d/*.dmb and ACMM_NESP/ACMM_NESP_model.ply, the
  cloud equal to the oracle's fusion of the written maps.
- Error paths: a hierarchy input that is not coarser, ACMM problems in the
  batched form.
"""
import ctypes as C
import os

import numpy as np
import pytest

from test_gpu_parity import depth_normal_agreement, make_problem, say

pytestmark = pytest.mark.gpu


def _acmm_params(hpm, **kw):
    p = hpm.default_params()
    p.acmm = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _same(tag, g_hyp, g_cost, o):
    ident = np.all(g_hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
    ident &= (g_cost.view(np.uint32) == o.cost.view(np.uint32))
    say(f"{tag}: identical pixels {ident.mean():.6f}")
    assert ident.all(), tag


@pytest.fixture(scope="module")
def small_scene(hpm):
    cfg = hpm.scene_config(seed=12, width=160, height=120, fx=125.0, fy=125.0, cx=80.0, cy=60.0)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    order = [2, 1, 3, 0, 4]
    return [cams[i] for i in order], [imgs[i].astype(np.float32) for i in order], gd[2]


def test_hierarchy_run_bitwise(hpm, cpu, small_scene):
    from oracle import pipeline

    fine, fimg, gt = small_scene
    coarse = [pipeline._scale_view(c, im, 80) for c, im in zip(fine, fimg)]
    cc, ci = [c for c, _ in coarse], [im for _, im in coarse]
    say("case: ACMM plain 80x60, then hierarchy RunPatchMatch at 160x120 through the C-ABI")
    o = cpu.Oracle(cc, ci, cpu.relaxed_params(cc, _acmm_params(hpm)), salt=0)
    o.run_patchmatch()
    pr = make_problem(hpm, cc, ci, _acmm_params(hpm))
    pr.RunPatchMatch()
    _same("coarse plain pass", pr.hypotheses(), pr.costs(), o)
    coarse_state = pr.hypotheses()
    pr.close()

    h = cpu.Oracle(fine, fimg, cpu.relaxed_params(fine, _acmm_params(hpm)), salt=0)
    h.set_hierarchy(o.hyp)
    h.run_patchmatch()
    pr = make_problem(hpm, fine, fimg, _acmm_params(hpm))
    cs = np.ascontiguousarray(coarse_state)
    rc = hpm.lib().hpm_set_hierarchy_input(pr._h, hpm.fptr(cs), cs.shape[1], cs.shape[0])
    assert rc == 0, hpm.lib().hpm_last_error(pr._h)
    pr.RunPatchMatch()
    _same("hierarchy pass", pr.hypotheses(), pr.costs(), h)
    acc = np.mean(np.abs(pr.hypotheses()[..., 3] - gt) / gt < 0.02)
    say(f"hierarchy pass accuracy (2%): {acc:.4f}")
    pr.close()


def test_geometric_pass_bitwise(hpm, cpu, small_scene):
    fine, fimg, gt = small_scene
    say("case: ACMM geometric pass (cap 3, weight 0.2, 2 iterations) at 160x120")
    o = cpu.Oracle(fine, fimg, cpu.relaxed_params(fine, _acmm_params(hpm)), salt=0)
    o.run_patchmatch()
    rng = np.random.default_rng(3)
    depths = [o.hyp[..., 3].copy()] + [gt * rng.uniform(0.97, 1.03, gt.shape).astype(np.float32)
                                       for _ in fine[1:]]
    p = _acmm_params(hpm, geom_consistency=1, max_iterations=2)
    g = cpu.Oracle(fine, fimg, cpu.relaxed_params(fine, p), depths=depths, salt=0)
    g.hyp[...] = o.hyp
    g.cost[...] = o.cost
    g.run_patchmatch()
    pr = hpm.Problem(0)
    pr.set_params(p)
    pr.input_init_memory(fine, fimg, depths=depths, ref_image_id=0)
    pr.CudaSpaceInitialization()
    pr.device_init_state(o.hyp, o.cost)
    pr.RunPatchMatch()
    assert pr.run_stats().launches == 4 + 2 * 2
    _same("geometric pass", pr.hypotheses(), pr.costs(), g)
    pr.close()


def test_acmm_schedule_matches_cpu(hpm):
    from oracle import pipeline
    from test_gpu_pipeline import scene

    cfg, cams, imgs, gd, problems = scene(hpm)
    opt = hpm.acmm_options(scale_bound=200, workers=4)
    say("case: ACMM schedule, 4 views 320x240, levels 160x120 + 320x240, 2 geometric passes each")
    pool, st = hpm.run_acmm(cams, imgs, problems, opt)
    say(f"gpu ACMM: runs={st.runs} n_eval={st.n_eval} device={st.ms_device:.1f} ms")
    assert st.runs == 2 * 4 * (1 + 2)
    cpu_state, d_first, d_geom, lcams = pipeline.run_cpu_acmm(cams, imgs, problems, opt)
    for v, _ in problems:
        assert pool.camera(v).width == 320 and pool.camera(v).height == 240
        g_hyp, g_cost = pool.state(v)
        c_hyp, c_cost = cpu_state[v]
        ok = depth_normal_agreement(g_hyp, c_hyp)
        ident = np.all(g_hyp.view(np.uint32) == c_hyp.view(np.uint32), axis=2).mean()
        gt = gd[v]
        acc = np.mean(np.abs(g_hyp[..., 3] - gt) / gt < 0.02)
        say(f"view {v}: agree={ok.mean():.5f} identical={ident:.5f} acc2%={acc:.4f}")
        assert ok.mean() >= 0.99
        assert np.array_equal(pool.depth(v, 1), d_geom[v]) or \
            np.mean(np.abs(pool.depth(v, 1) - d_geom[v]) <= 0.01 * np.abs(d_geom[v])) >= 0.99
    # ACMM's RunFusion on the last level, device records + host scan against
    # the oracle's sequential fusion of the same maps
    from oracle import cpu

    pts = hpm.fusion(pool, problems, geom_consistency=True, acmm=True)
    colors = [np.repeat(np.clip(im, 0, 255).astype(np.uint8)[..., None], 3, axis=2) for im in imgs]
    want = cpu.fusion([pool.camera(v) for v in range(4)], [pool.depth(v, 1) for v in range(4)],
                      [pool.state(v)[0][..., :3].copy() for v in range(4)], colors, problems, hpm.POINT_DTYPE,
                      acmm=True)
    say(f"ACMM fusion: {len(pts)} points (oracle {len(want)})")
    assert len(pts) == len(want) > 0
    assert np.array_equal(pts.view(np.uint8), want.view(np.uint8))
    pool.close()


@pytest.mark.parametrize("mode", [[], ["--class-api"], ["--scale-bound", "100"],
                                  ["--class-api", "--scale-bound", "100"]])
def test_executable_acmm_on_folder(hpm, tmp_path, mode):
    """The device schedule and the reference's main written against the
    drop-in `hpm::ACMM` class, at one level and at two (scale bound 100:
    96x72 then 192x144, the second level started by the hierarchy pass)."""
    import subprocess

    from oracle import cpu

    exe = os.path.join(os.path.dirname(hpm.LIB_PATH), "hpm_mvs")
    cfg = hpm.scene_config(seed=6, width=192, height=144, fx=150.0, fy=150.0, cx=96.0, cy=72.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    say(f"case: hpm_mvs --acmm {' '.join(mode)} on a 3-view 192x144 folder")
    r = subprocess.run([exe, str(tmp_path), "--acmm"] + mode, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    say(r.stdout.strip().splitlines()[-1])
    cams, depths, normals, colors = [], [], [], []
    for v in range(3):
        base = tmp_path / "ACMM_NESP" / f"2333_{v:08d}"
        d = hpm.read_dmb(str(base / "depths_geom.dmb"))
        gt = hpm.read_dmb(str(tmp_path / "gt" / f"{v:08d}_depth.dmb"))
        acc = np.mean(np.abs(d - gt) / gt < 0.05)
        say(f"view {v}: within 5% of the true depth {acc:.4f}")
        assert d.shape == (144, 192) and acc > 0.3
        for f in ("depths.dmb", "normals.dmb", "costs.dmb"):
            assert (base / f).exists()
        cam = hpm.read_camera(str(tmp_path / "cams" / f"{v:08d}_cam.txt"))
        img = hpm.read_pgm(str(tmp_path / "images" / f"{v:08d}.pgm"))
        cam.height, cam.width = img.shape
        cams.append(cam)
        depths.append(d)
        normals.append(hpm.read_dmb(str(base / "normals.dmb")))
        colors.append(np.repeat(np.clip(img, 0, 255).astype(np.uint8)[..., None], 3, axis=2))
    ply = (tmp_path / "ACMM_NESP" / "ACMM_NESP_model.ply").read_bytes()
    head, _, body = ply.partition(b"end_header\n")
    n = int([ln for ln in head.split(b"\n") if ln.startswith(b"element vertex")][0].split()[-1])
    rows = np.frombuffer(body, dtype=np.dtype([("xyz", "<f4", 3), ("n", "<f4", 3), ("rgb", "u1", 3)]))
    problems = hpm.read_pair(str(tmp_path / "pair.txt"))
    want = cpu.fusion(cams, depths, normals, colors, [(r_, list(s)) for r_, s in problems], hpm.POINT_DTYPE,
                      acmm=True)
    say(f"fused points: {n} (oracle {len(want)})")
    assert n == len(want) and n > 0
    assert np.array_equal(rows["xyz"], np.stack([want["x"], want["y"], want["z"]], axis=1))
    assert np.array_equal(rows["n"], np.stack([want["nx"], want["ny"], want["nz"]], axis=1))
    assert np.array_equal(rows["rgb"], np.stack([want["r"], want["g"], want["b"]], axis=1))


def test_class_api_replays_device_schedule(hpm, tmp_path):
    """`hpm_mvs --acmm --class-api` (the reference's main over `hpm::ACMM`,
    .dmb files between passes) with every pass on the Philox stream the
    device schedule gives it (hpm_set_stream_index): every map file and the
    fused cloud equal the device schedule's byte for byte, over two levels
    (96x72, then 192x144 started by the hierarchy pass)."""
    import shutil
    import subprocess

    exe = os.path.join(os.path.dirname(hpm.LIB_PATH), "hpm_mvs")
    cfg = hpm.scene_config(seed=8, width=192, height=144, fx=150.0, fy=150.0, cx=96.0, cy=72.0, num_views=3)
    dev, cls = tmp_path / "device", tmp_path / "class"
    hpm.write_scene_folder(cfg, str(dev))
    shutil.copytree(dev, cls)
    say("case: hpm_mvs --acmm --scale-bound 100, device schedule and class API, 3 views 192x144")
    for folder, mode in ((dev, []), (cls, ["--class-api"])):
        r = subprocess.run([exe, str(folder), "--acmm", "--scale-bound", "100"] + mode, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
    for v in range(3):
        for f in ("depths.dmb", "depths_geom.dmb", "normals.dmb", "costs.dmb"):
            x = (dev / "ACMM_NESP" / f"2333_{v:08d}" / f).read_bytes()
            y = (cls / "ACMM_NESP" / f"2333_{v:08d}" / f).read_bytes()
            assert x == y, (v, f)
    x = (dev / "ACMM_NESP" / "ACMM_NESP_model.ply").read_bytes()
    y = (cls / "ACMM_NESP" / "ACMM_NESP_model.ply").read_bytes()
    say(f"maps identical; PLY {len(x)} bytes, identical {x == y}")
    assert x == y


def test_acmm_error_paths(hpm, small_scene):
    fine, fimg, gt = small_scene
    say("case: ACMM error paths")
    pr = make_problem(hpm, fine, fimg, _acmm_params(hpm))
    same_size = np.zeros((120, 160, 4), np.float32)
    assert hpm.lib().hpm_set_hierarchy_input(pr._h, hpm.fptr(same_size), 160, 120) != 0
    assert b"coarser" in hpm.lib().hpm_last_error(pr._h)
    pr.close()
    pr = make_problem(hpm, fine, fimg, _acmm_params(hpm))
    batch = hpm.Batch(0)
    with pytest.raises(hpm.HpmError):
        batch.run([pr])
    batch.close()
    pr.close()
    bad = _acmm_params(hpm, max_iterations=4)
    p2 = hpm.Problem(0)
    with pytest.raises(hpm.HpmError):
        p2.set_params(bad)
    p2.close()
