"""The device pass schedule (hpm_pipeline_run: plain pass, 0.5x prior with
JBU, 1x prior with the KNN supplement, geometric-consistency passes) against
its CPU restatement (oracle/pipeline.py) on a 4-view 320x240 scene, with the
scale bound lowered to 200 so that the JBU level runs: the kernels are
bit-identical stage by stage, so hypotheses, costs and depth maps must match
bit for bit."""
import os

import numpy as np
import pytest

from test_gpu_parity import depth_normal_agreement, say

pytestmark = pytest.mark.gpu


def scene(hpm):
    cfg = hpm.scene_config(seed=21, width=320, height=240, fx=250.0, fy=250.0, cx=160.0, cy=120.0, num_views=4,
                           num_planes=4, num_boxes=6, low_texture_fraction=0.3)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    n = cfg.num_views
    problems = [(r, sorted([v for v in range(n) if v != r], key=lambda v: (abs(v - r), v))) for r in range(n)]
    return cfg, cams, [im.astype(np.float32) for im in imgs], gd, problems


@pytest.mark.parametrize("knn_k,geom", [(6, 2), (8, 3)])
def test_pipeline_matches_cpu_schedule(hpm, knn_k, geom):
    """The reference's settings (KNN K = 6, 2 geometric passes) and the
    1600x1066 workload's (K = 8, 3 passes)."""
    from oracle import pipeline

    cfg, cams, imgs, gd, problems = scene(hpm)
    opt = hpm.pipeline_options(scale_bound=200)
    opt.params.geom_iterations = geom
    opt.params.knn_k = knn_k
    say(f"case: pipeline 4 views 320x240, prior scales 0.5x+1x, K={knn_k}, {geom} geometric passes")
    pool = hpm.Pool(cams, imgs)
    st = hpm.run_pipeline(pool, problems, opt)
    say(f"gpu pipeline: runs={st.runs} n_eval={st.n_eval} wall={st.ms_wall:.1f} ms device={st.ms_device:.1f} ms "
        f"prior host={st.ms_prior_host:.1f} ms")
    assert st.runs == 4 * (3 + geom)
    cpu_state, d_plain, d_geom = pipeline.run_cpu_pipeline(cams, imgs, problems, opt)
    for v, _ in problems:
        g_hyp, g_cost = pool.state(v)
        c_hyp, c_cost = cpu_state[v]
        ok = depth_normal_agreement(g_hyp, c_hyp)
        ident = np.all(g_hyp.view(np.uint32) == c_hyp.view(np.uint32), axis=2).mean()
        same_c = (g_cost.view(np.uint32) == c_cost.view(np.uint32)) | (np.isnan(g_cost) & np.isnan(c_cost))
        gt = gd[v]
        acc = np.mean(np.abs(g_hyp[..., 3] - gt) / gt < 0.02)
        say(f"view {v}: agree={ok.mean():.5f} identical={ident:.5f} costs={same_c.mean():.5f} acc2%={acc:.4f}")
        assert ident == 1.0 and same_c.all()
        assert np.array_equal(pool.depth(v, 0), d_plain[v])
        assert np.array_equal(pool.depth(v, 1), d_geom[v])
    pool.close()


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "hpm-mvs_amd", "unet_prior.bin")


@pytest.mark.parametrize("mode", [[], ["--class-api"], ["--learned-prior", WEIGHTS]])
def test_executable_on_folder(hpm, tmp_path, mode):
    """The `HPM <dense_folder>` executable: device schedule, the reference's
    ProcessProblem written against the drop-in C++ class, and the device
    schedule with the learned prior."""
    import subprocess

    exe = os.path.join(os.path.dirname(hpm.LIB_PATH), "hpm_mvs")
    cfg = hpm.scene_config(seed=6, width=192, height=144, fx=150.0, fy=150.0, cx=96.0, cy=72.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    say(f"case: hpm_mvs {' '.join(mode)} on a 3-view 192x144 folder")
    r = subprocess.run([exe, str(tmp_path), "--geom-iterations", "1"] + mode, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    for v in range(3):
        base = tmp_path / "HPM" / f"2333_{v:08d}"
        d = hpm.read_dmb(str(base / "depths_geom.dmb"))
        gt = hpm.read_dmb(str(tmp_path / "gt" / f"{v:08d}_depth.dmb"))
        assert np.mean(np.abs(d - gt) / gt < 0.05) > 0.3
        assert hpm.read_dmb(str(base / "costs.dmb")).shape == (144, 192)
    # RunFusion (main.cpp:675) after the passes: the fused cloud equals the
    # oracle's fusion of the maps the run wrote
    from oracle import cpu

    ply = (tmp_path / "HPM" / "HPM_model.ply").read_bytes()
    head, _, body = ply.partition(b"end_header\n")
    n = int([ln for ln in head.split(b"\n") if ln.startswith(b"element vertex")][0].split()[-1])
    assert len(body) == 27 * n
    rows = np.frombuffer(body, dtype=np.dtype([("xyz", "<f4", 3), ("n", "<f4", 3), ("rgb", "u1", 3)]))
    problems = hpm.read_pair(str(tmp_path / "pair.txt"))
    cams, depths, normals, colors = [], [], [], []
    for v in range(3):
        cam = hpm.read_camera(str(tmp_path / "cams" / f"{v:08d}_cam.txt"))
        img = hpm.read_pgm(str(tmp_path / "images" / f"{v:08d}.pgm"))
        cam.height, cam.width = img.shape
        base = tmp_path / "HPM" / f"2333_{v:08d}"
        cams.append(cam)
        depths.append(hpm.read_dmb(str(base / "depths_geom.dmb")))
        normals.append(hpm.read_dmb(str(base / "normals.dmb")))
        colors.append(np.repeat(np.clip(img, 0, 255).astype(np.uint8)[..., None], 3, axis=2))
    want = cpu.fusion(cams, depths, normals, colors, [(r, list(s)) for r, s in problems], hpm.POINT_DTYPE)
    say(f"fused points: {n} (oracle {len(want)})")
    assert n == len(want) and n > 0
    assert np.array_equal(rows["xyz"], np.stack([want["x"], want["y"], want["z"]], axis=1))
    assert np.array_equal(rows["n"], np.stack([want["nx"], want["ny"], want["nz"]], axis=1))
    assert np.array_equal(rows["rgb"], np.stack([want["r"], want["g"], want["b"]], axis=1))


def test_class_api_replays_device_schedule(hpm, tmp_path):
    """`hpm_mvs --class-api` (the reference's ProcessProblem over `hpm::HPM`,
    host prior construction, .dmb files between passes) with every pass on
    the Philox stream the device schedule gives it: every map file and the
    fused cloud equal the device schedule's (device prior) byte for byte."""
    import shutil
    import subprocess

    exe = os.path.join(os.path.dirname(hpm.LIB_PATH), "hpm_mvs")
    cfg = hpm.scene_config(seed=6, width=192, height=144, fx=150.0, fy=150.0, cx=96.0, cy=72.0, num_views=3)
    dev, cls = tmp_path / "device", tmp_path / "class"
    hpm.write_scene_folder(cfg, str(dev))
    shutil.copytree(dev, cls)
    say("case: hpm_mvs --geom-iterations 2, device schedule and class API, 3 views 192x144")
    for folder, mode in ((dev, []), (cls, ["--class-api"])):
        r = subprocess.run([exe, str(folder), "--geom-iterations", "2"] + mode, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
    for v in range(3):
        for f in ("depths.dmb", "depths_geom.dmb", "normals.dmb", "costs.dmb"):
            x = hpm.read_dmb(str(dev / "HPM" / f"2333_{v:08d}" / f))
            y = hpm.read_dmb(str(cls / "HPM" / f"2333_{v:08d}" / f))
            same = np.mean(x.view(np.uint32) == y.view(np.uint32))
            say(f"view {v} {f}: identical {same:.6f}")
            assert same == 1.0, (v, f)
    x = (dev / "HPM" / "HPM_model.ply").read_bytes()
    y = (cls / "HPM" / "HPM_model.ply").read_bytes()
    say(f"PLY {len(x)} bytes, identical {x == y}")
    assert x == y


def test_pipeline_folder_round_trip(hpm, tmp_path):
    cfg = hpm.scene_config(seed=4, width=160, height=120, fx=125.0, fy=125.0, cx=80.0, cy=60.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    opt = hpm.pipeline_options()
    opt.params.geom_iterations = 1
    say("case: pipeline on a 3-view 160x120 folder")
    st = hpm.run_pipeline_folder(str(tmp_path), opt)
    assert st.runs == 3 * (2 + 1)  # k = 0: plain + 1x prior per view, then one geometric pass
    for v in range(3):
        d = hpm.read_dmb(str(tmp_path / "HPM" / f"2333_{v:08d}" / "depths_geom.dmb"))
        n = hpm.read_dmb(str(tmp_path / "HPM" / f"2333_{v:08d}" / "normals.dmb"))
        assert d.shape == (120, 160) and n.shape == (120, 160, 3)
        gt = hpm.read_dmb(str(tmp_path / "gt" / f"{v:08d}_depth.dmb"))
        assert np.mean(np.abs(d - gt) / gt < 0.05) > 0.3


def test_executable_on_jpeg_folder(hpm, tmp_path):
    """The reference's folder layout with colour images/%08d.jpg (as real
    datasets ship them): hpm_mvs decodes the JPEGs (luminance for the
    matching, IMREAD_GRAYSCALE; colour for the fused points,
    IMREAD_COLOR, main.cpp:523) and reconstructs; the PLY equals the
    oracle's fusion of the written maps with the decoded colours."""
    import subprocess

    Image = pytest.importorskip("PIL.Image")
    exe = os.path.join(os.path.dirname(hpm.LIB_PATH), "hpm_mvs")
    cfg = hpm.scene_config(seed=6, width=192, height=144, fx=150.0, fy=150.0, cx=96.0, cy=72.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    for v in range(3):
        pgm = tmp_path / "images" / f"{v:08d}.pgm"
        img = hpm.read_image_grey(str(pgm)).astype(np.int32)
        rgb = np.stack([img + 25, img, img - 25 + (np.arange(img.shape[1]) % 7)], -1)
        Image.fromarray(np.clip(rgb, 0, 255).astype(np.uint8)).save(str(tmp_path / "images" / f"{v:08d}.jpg"),
                                                                   "JPEG", quality=95)
        pgm.unlink()
    say("case: hpm_mvs on a 3-view 192x144 folder of JPEG images")
    r = subprocess.run([exe, str(tmp_path), "--geom-iterations", "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    accs = []
    for v in range(3):
        d = hpm.read_dmb(str(tmp_path / "HPM" / f"2333_{v:08d}" / "depths_geom.dmb"))
        gt = hpm.read_dmb(str(tmp_path / "gt" / f"{v:08d}_depth.dmb"))
        accs.append(float(np.mean(np.abs(d - gt) / gt < 0.05)))
    say(f"accuracy (5%) per view: {accs}")
    assert min(accs) > 0.3
    from oracle import cpu

    ply = (tmp_path / "HPM" / "HPM_model.ply").read_bytes()
    head, _, body = ply.partition(b"end_header\n")
    n = int([ln for ln in head.split(b"\n") if ln.startswith(b"element vertex")][0].split()[-1])
    rows = np.frombuffer(body, dtype=np.dtype([("xyz", "<f4", 3), ("n", "<f4", 3), ("rgb", "u1", 3)]))
    problems = hpm.read_pair(str(tmp_path / "pair.txt"))
    cams, depths, normals, colors = [], [], [], []
    for v in range(3):
        cam = hpm.read_camera(str(tmp_path / "cams" / f"{v:08d}_cam.txt"))
        col = hpm.read_image_rgb(str(tmp_path / "images" / f"{v:08d}.jpg"))
        cam.height, cam.width = col.shape[:2]
        base = tmp_path / "HPM" / f"2333_{v:08d}"
        cams.append(cam)
        depths.append(hpm.read_dmb(str(base / "depths_geom.dmb")))
        normals.append(hpm.read_dmb(str(base / "normals.dmb")))
        colors.append(col)
    want = cpu.fusion(cams, depths, normals, colors, [(r, list(s)) for r, s in problems], hpm.POINT_DTYPE)
    say(f"fused points: {n} (oracle {len(want)}), distinct colours {len(np.unique(rows['rgb'], axis=0))}")
    assert n == len(want) and n > 0
    assert np.array_equal(rows["rgb"], np.stack([want["r"], want["g"], want["b"]], axis=1))
    assert np.mean(rows["rgb"][:, 0].astype(int) - rows["rgb"][:, 2].astype(int)) > 30  # red above blue


@pytest.mark.parametrize("size", [(320, 240), (1600, 1066)])
def test_device_prior_matches_host_construction(hpm, size):
    """hpm_build_planar_prior (support points, their Delaunay triangulation,
    rasterisation, triangle planes, KNN grid and supplement and assembly, all
    on the device) against
    the host construction of the same scale (oracle/pipeline.py:
    prior_scale_maps, the library's host functions), at 0.5x (through the
    JBU) and 1x, bit for bit."""
    import ctypes as C

    from oracle import pipeline
    from test_gpu_parity import make_problem

    w, h = size
    f = 0.78 * w
    cfg = hpm.scene_config(seed=31, width=w, height=h, fx=f, fy=f, cx=w / 2.0, cy=h / 2.0, num_views=4,
                           num_planes=20, num_boxes=50, low_texture_fraction=0.3)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg, want_gt=False)
    params = hpm.default_params()
    params.knn_k = 8
    pr = make_problem(hpm, cams, [im.astype(np.float32) for im in imgs], params)
    pr.RunPatchMatch()
    hyp, cost = pr.hypotheses(), pr.costs()
    p = pr.params()
    for factor in (0.5, 1.0):
        say(f"case: device planar prior {w}x{h} at {factor}x")
        host_ms = C.c_double(0.0)
        rc = hpm.lib().hpm_build_planar_prior(pr._h, factor, 8, C.byref(host_ms))
        assert rc == 0, hpm.lib().hpm_last_error(pr._h)
        planes = np.empty((h, w, 4), np.float32)
        mask = np.empty((h, w), np.uint32)
        assert hpm.lib().hpm_debug_read_prior(pr._h, hpm.fptr(planes), hpm.uptr(mask)) == 0
        want_p, want_m = pipeline.prior_scale_maps(cams[0], np.ascontiguousarray(hyp[..., 3]), cost,
                                                   imgs[0].astype(np.float32), factor, 8, np.float32(p.depth_min),
                                                   np.float32(p.depth_max))
        same = np.all(planes.view(np.uint32) == want_p.view(np.uint32), axis=2) & (mask == want_m)
        say(f"{factor}x: mask on {mask.mean():.4f}, identical pixels {same.mean():.6f}, host part "
            f"{host_ms.value:.1f} ms")
        assert mask.mean() > 0.3
        assert same.all()
    pr.close()


def test_pipeline_device_prior_equals_host_prior(hpm):
    """The pass schedule with the planar prior built on the device and on the
    host: the same maps"""
    cfg, cams, imgs, gd, problems = scene(hpm)
    say("case: pipeline with the device prior and with the host prior, 4 views 320x240")
    out = []
    for dev in (1, 0):
        opt = hpm.pipeline_options(scale_bound=200, prior_on_device=dev)
        opt.params.geom_iterations = 1
        pool = hpm.Pool(cams, imgs)
        st = hpm.run_pipeline(pool, problems, opt)
        say(f"prior_on_device={dev}: prior host {st.ms_prior_host:.1f} ms, device {st.ms_prior_device:.1f} ms")
        out.append([pool.state(v) for v, _ in problems])
        pool.close()
    for (a_h, a_c), (b_h, b_c) in zip(*out):
        assert np.array_equal(a_h.view(np.uint32), b_h.view(np.uint32))
        assert np.array_equal(a_c.view(np.uint32), b_c.view(np.uint32))


@pytest.mark.parametrize("case", ["textureless", "tiny"])
def test_device_prior_degenerate(hpm, case):
    """The device planar prior where the host construction degenerates: a
    textureless reference (every cost 2, no support point, no triangle, no
    KNN plane) and a 40x30 view (a handful of support points): both
    constructions agree, also on an empty prior."""
    import ctypes as C

    from oracle import pipeline
    from test_gpu_parity import make_problem

    w, h = (160, 120) if case == "textureless" else (40, 30)
    cfg = hpm.scene_config(seed=33, width=w, height=h, fx=0.78 * w, fy=0.78 * w, cx=w / 2.0, cy=h / 2.0,
                           num_views=3, num_planes=4, num_boxes=2)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg, want_gt=False)
    imgs = [im.astype(np.float32) for im in imgs]
    if case == "textureless":
        imgs = [np.full_like(im, 128.0) for im in imgs]
    params = hpm.default_params()
    params.knn_k = 6
    pr = make_problem(hpm, cams, imgs, params)
    pr.RunPatchMatch()
    hyp, cost = pr.hypotheses(), pr.costs()
    p = pr.params()
    for factor in (0.5, 1.0):
        say(f"case: device planar prior, {case} {w}x{h}, {factor}x")
        rc = hpm.lib().hpm_build_planar_prior(pr._h, factor, 6, None)
        assert rc == 0, hpm.lib().hpm_last_error(pr._h)
        planes = np.empty((h, w, 4), np.float32)
        mask = np.empty((h, w), np.uint32)
        assert hpm.lib().hpm_debug_read_prior(pr._h, hpm.fptr(planes), hpm.uptr(mask)) == 0
        want_p, want_m = pipeline.prior_scale_maps(cams[0], np.ascontiguousarray(hyp[..., 3]), cost, imgs[0], factor, 6,
                                                   np.float32(p.depth_min), np.float32(p.depth_max))
        say(f"mask on {mask.mean():.4f} (host {want_m.mean():.4f})")
        assert np.array_equal(mask, want_m)
        assert np.array_equal(planes.view(np.uint32), want_p.view(np.uint32))
        if case == "textureless":
            assert mask.sum() == 0
    pr.close()


@pytest.mark.gpu
def test_device_init_rejects_a_mis_sized_reload(hpm, tmp_path):
    """With geometric consistency, CudaSpaceInitialization reloads the view's
    depths, normals and costs (HPM.cpp:662-699). A normals.dmb of another
    size (an earlier run at another scale) is an I/O error, not a read past
    the buffers; the files at the view's size load."""
    W, H = 96, 64
    cfg = hpm.scene_config(seed=6, width=W, height=H, fx=80.0, fy=80.0, cx=48.0, cy=32.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    for v in range(3):
        d = tmp_path / "HPM" / f"2333_{v:08d}"
        d.mkdir(parents=True, exist_ok=True)
        hpm.write_dmb(str(d / "depths.dmb"), np.full((H, W), 4.0, np.float32))
        nrm = np.zeros((H, W, 3), np.float32)
        nrm[..., 2] = -1.0
        hpm.write_dmb(str(d / "normals.dmb"), nrm)
        hpm.write_dmb(str(d / "costs.dmb"), np.full((H, W), 0.5, np.float32))
    big = np.zeros((2 * H, 2 * W, 3), np.float32)
    hpm.write_dmb(str(tmp_path / "HPM" / "2333_00000000" / "normals.dmb"), big)
    say("case: CudaSpaceInitialization with a 192x128 normals.dmb for a 96x64 view")
    pr = hpm.Problem()
    pr.SetGeomConsistencyParams(False)
    pr.InuputInitialization(str(tmp_path), 0, [1, 2])
    with pytest.raises(hpm.HpmError, match="not 96x64"):
        pr.CudaSpaceInitialization(str(tmp_path))
    pr.close()
    say("case: the same with the files at the view's size")
    hpm.write_dmb(str(tmp_path / "HPM" / "2333_00000000" / "normals.dmb"), big[:H, :W])
    pr = hpm.Problem()
    pr.SetGeomConsistencyParams(False)
    pr.InuputInitialization(str(tmp_path), 0, [1, 2])
    pr.CudaSpaceInitialization(str(tmp_path))
    pr.RunPatchMatch()
    assert pr.hypotheses().shape == (H, W, 4)
    pr.close()


@pytest.mark.parametrize("learned", [False, True])
def test_batched_schedule_equals_per_view_schedule(hpm, learned):
    """hpm_pipeline_options.batch: the plain pass, the learned-prior pass and
    the first geometric pass of all views as one hpm_run_batch call each
    give the per-view schedule's maps bit for bit, on the first run (tuned
    and captured) and on a second run (the captured schedules replayed). The
    per-view schedule runs twice on its pool, so its maps come from handles
    (and a learned prior) the first run left in the pool, the batched form's
    first run from new ones."""
    cfg, cams, imgs, gd, problems = scene(hpm)

    def options(batch=None):
        opt = hpm.pipeline_options(scale_bound=200)
        opt.params.geom_iterations = 2
        if learned:
            opt.learned_prior = os.path.abspath(WEIGHTS).encode()
        if batch is not None:
            opt.batch = batch._b
        return opt

    say(f"case: batched schedule, 4 views 320x240, learned prior {learned}")
    ref_pool = hpm.Pool(cams, imgs)
    for _ in range(2):
        hpm.run_pipeline(ref_pool, problems, options())
    batch = hpm.Batch(0)
    pool = hpm.Pool(cams, imgs)
    for run in range(2):
        st = hpm.run_pipeline(pool, problems, options(batch))
        say(f"run {run}: {st.batches} batches, {st.runs} RunPatchMatch, learned views {st.learned_views}")
        assert st.batches == 2 + (1 if st.learned_views else 0)
        assert st.learned_views == (4 if learned else 0) or st.learned_fallbacks > 0
        for v, _ in problems:
            a, b = ref_pool.state(v), pool.state(v)
            assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), (run, v)
            assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), (run, v)
            assert np.array_equal(ref_pool.depth(v, 0), pool.depth(v, 0))
            assert np.array_equal(ref_pool.depth(v, 1), pool.depth(v, 1))
    say(batch.log().strip())
    batch.close()
    pool.close()
    ref_pool.close()
