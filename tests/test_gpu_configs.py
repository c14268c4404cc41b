"""The batched and large-view workloads against the CPU oracle, and the
max_image_size downscale of InuputInitialization.

  - 32 references of uneven size with 4-10 sources each (the batched
    workload at reduced sizes) through one hpm_run_batch call: every
    problem bit-identical to the oracle's RunPatchMatch, on the first call
    and on the replay of the captured schedule;
  - one 3840x2160 reference with 20 sources at full resolution: the
    initialisation over the whole view and one red/black half-iteration in
    three row bands (the top and bottom borders, where the last views' pool
    offsets are largest, and the middle) bit-identical to the oracle; a full
    RunPatchMatch finite (but for the NaN costs the reference's fallback
    leaves), costs in [0, 2], the evaluation count inside its closed-form
    bracket, and the depth accuracy against ground truth;
  - views above max_image_size are resized and their cameras scaled
    (HPM.cpp:524-551) before anything else: the run equals the oracle's on
    the resized images and scaled cameras, bit for bit;
  - a pool above 2 GiB (its entries are addressed by unsigned 32-bit byte
    offsets): a problem whose views lie past 2^31 bytes gives the same maps
    as the same views at small offsets.
"""
import numpy as np
import pytest

from conftest import problem_inputs
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _same(g_hyp, g_cost, o):
    sh = np.all(g_hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
    sc = (g_cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(g_cost) & np.isnan(o.cost))
    return sh & sc


def test_batch_of_32_uneven_references_matches_oracle(hpm, cpu):
    rng = np.random.default_rng(3)
    problems, oracles = [], []
    shapes = []
    for i in range(32):
        w = int(rng.integers(64, 201))
        h = int(rng.integers(24, 76)) * 2
        nsrc = int(rng.integers(4, 11))
        f = 0.8 * w
        cfg = hpm.scene_config(seed=3 + i, width=w, height=h, fx=f, fy=f, cx=w / 2.0, cy=h / 2.0,
                               num_views=nsrc + 1, num_planes=6, num_boxes=8, low_texture_fraction=0.3)
        cams, imgs, gd, gn, low, order = problem_inputs(hpm, cfg, (nsrc + 1) // 2)
        pr = hpm.Problem(0)
        pr.input_init_memory(cams, imgs, ref_image_id=i)
        pr.CudaSpaceInitialization()
        problems.append(pr)
        oracles.append(cpu.Oracle(cams, imgs, pr.params(), salt=i))
        shapes.append((w, h, nsrc))
    say(f"case: batch of 32 references, sizes {min(s[0] for s in shapes)}..{max(s[0] for s in shapes)} x "
        f"{min(s[1] for s in shapes)}..{max(s[1] for s in shapes)}, 4-10 sources")
    batch = hpm.Batch(0)
    for call in range(2):
        st = batch.run(problems)
        say(f"call {call}: grid {st.grid}, reused {st.graph_reused}, {st.ms_device:.2f} ms device, "
            f"{st.n_eval} evaluations")
        bad = []
        for i, (pr, o) in enumerate(zip(problems, oracles)):
            o.run_patchmatch()
            same = _same(pr.hypotheses(), pr.costs(), o)
            if not same.all() or pr.run_stats().n_eval != o.n_eval:
                bad.append((i, shapes[i], float(same.mean()), pr.run_stats().n_eval, o.n_eval))
        say(f"call {call}: problems differing from the oracle: {bad}")
        assert not bad
    batch.close()
    for pr in problems:
        pr.close()


@pytest.fixture(scope="module")
def c5_problem(hpm):
    cfg = hpm.workload_config("c5")
    say("rendering the 21-view 3840x2160 scene")
    cams, imgs, gd, gn, low, order = problem_inputs(hpm, cfg, 10)
    return cams, imgs, gd[order[0]]


def test_large_view_stages_match_oracle(hpm, cpu, c5_problem):
    cams, imgs, gt = c5_problem
    say("case: 3840x2160, 20 sources: init and one half-iteration against the oracle")
    p = hpm.default_params()
    p.max_image_size = 4096
    pr = hpm.Problem(0)
    pr.set_params(p)
    pr.input_init_memory(cams, imgs, ref_image_id=0)
    pr.CudaSpaceInitialization()
    W, H = pr.dims()
    assert (W, H) == (3840, 2160)
    o = cpu.Oracle(cams, imgs, pr.params(), salt=0)
    n_init = o.init()
    pr.debug_stage(0)
    hyp, cost, sel = pr.debug_state()
    same = _same(hyp, cost, o) & (sel == o.sel)
    say(f"init: identical pixels {same.mean():.7f}, {n_init} evaluations")
    assert same.all()
    bands = [(0, 48), (1056, 1104), (2112, 2160)]
    for y0, y1 in bands:
        o.half_iteration_rows(0, 0, y0, y1)
    pr.debug_stage(1, 0, 0)
    hyp, cost, sel = pr.debug_state()
    for y0, y1 in bands:
        g = hyp[y0:y1], cost[y0:y1], sel[y0:y1]
        s = (np.all(g[0].view(np.uint32) == o.hyp[y0:y1].view(np.uint32), axis=2)
             & ((g[1].view(np.uint32) == o.cost[y0:y1].view(np.uint32)) | (np.isnan(g[1]) & np.isnan(o.cost[y0:y1])))
             & (g[2] == o.sel[y0:y1]))
        say(f"half-iteration rows {y0}..{y1}: identical pixels {s.mean():.7f}")
        assert s.all(), (y0, y1)
    pr.close()


def test_large_view_full_run_properties(hpm, c5_problem):
    cams, imgs, gt = c5_problem
    say("case: 3840x2160, 20 sources: full RunPatchMatch")
    p = hpm.default_params()
    p.max_image_size = 4096
    pr = hpm.Problem(0)
    pr.set_params(p)
    pr.input_init_memory(cams, imgs, ref_image_id=0)
    pr.CudaSpaceInitialization()
    pr.RunPatchMatch()
    hyp, cost = pr.hypotheses(), pr.costs()
    st = pr.run_stats()
    W, H, N1, I = 3840, 2160, 20, 3
    acc = float(np.mean(np.abs(hyp[..., 3] - gt) / gt < 0.02))
    say(f"n_eval {st.n_eval} ({st.n_eval / (N1 * W * H):.2f} per pixel and view), {st.ms_total:.1f} ms, "
        f"accuracy (2%) {acc:.4f}")
    # a pixel whose sampling probabilities are all zero keeps a NaN cost
    # (the non-prior fallback assigns cost_now, HPM.cu:1716-1721)
    nan = np.isnan(cost)
    say(f"NaN costs: {int(nan.sum())}")
    assert np.all(np.isfinite(hyp)) and nan.mean() < 1e-4
    assert cost[~nan].min() >= 0.0 and cost[~nan].max() <= 2.0
    # init: one evaluation per (pixel, view); per iteration 8 neighbour
    # directions on every view (fewer near the border), up to 3 extension
    # rounds, the re-score and 5 refinement candidates on the drawn views
    assert 0.97 * N1 * W * H * (1 + I * 8) <= st.n_eval <= N1 * W * H * (1 + I * 17)
    assert acc > 0.9
    pr.close()


def test_downscale_above_max_image_size(hpm, cpu):
    """A 200x150 problem with max_image_size 160: every view becomes 160x120
    (factor min(160/200, 160/150) = 0.8), K scaled by 160/200 and 120/150,
    and the run equals the oracle's on those images and cameras."""
    import copy

    w, h, n = 200, 150, 5
    cfg = hpm.scene_config(seed=9, width=w, height=h, fx=160.0, fy=160.0, cx=100.0, cy=75.0, num_views=n)
    cams, imgs, gd, gn, low, order = problem_inputs(hpm, cfg, 2)
    p = hpm.default_params()
    p.max_image_size = 160
    pr = hpm.Problem(0)
    pr.set_params(p)
    pr.input_init_memory(cams, imgs, ref_image_id=0)
    pr.CudaSpaceInitialization()
    assert pr.dims() == (160, 120)
    scams, simgs = [], []
    for c, im in zip(cams, imgs):
        s = copy.copy(c)
        s.K[0] = float(np.float32(c.K[0]) * np.float32(160 / np.float32(200)))
        s.K[2] = float(np.float32(c.K[2]) * np.float32(160 / np.float32(200)))
        s.K[4] = float(np.float32(c.K[4]) * np.float32(120 / np.float32(150)))
        s.K[5] = float(np.float32(c.K[5]) * np.float32(120 / np.float32(150)))
        s.width, s.height = 160, 120
        scams.append(s)
        simgs.append(hpm.resize_bilinear(im, 160, 120))
    for i in range(n):
        got = pr.camera(i)
        assert list(got.K) == list(scams[i].K), i
    assert any(not np.array_equal(im, np.round(im)) for im in simgs)  # resized grey values are fractional
    say("case: downscaled 200x150 -> 160x120, 4 sources")
    pr.RunPatchMatch()
    o = cpu.Oracle(scams, simgs, pr.params(), salt=0)
    o.run_patchmatch()
    same = _same(pr.hypotheses(), pr.costs(), o)
    say(f"identical pixels {same.mean():.6f}, n_eval gpu={pr.run_stats().n_eval} cpu={o.n_eval}")
    assert same.all() and pr.run_stats().n_eval == o.n_eval
    pr.close()


def test_pool_beyond_2gib(hpm):
    """38 views: 35 fillers of 3200x2400 put the last three views' footprint
    entries (8 bytes each) past byte 2^31 of the pool; the same three views
    behind 35 fillers of 16x16 sit at small offsets. The pass schedule's
    plain RunPatchMatch of reference 37 with sources 36 and 35 (same ids, so
    the same random streams) must give bit-identical maps in both pools."""
    import copy

    W, H = 3200, 2400
    cfg = hpm.scene_config(seed=31, width=W, height=H, fx=2500.0, fy=2500.0, cx=1600.0, cy=1200.0, num_views=3)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg, want_gt=False)
    imgs = [im.astype(np.float32) for im in imgs]
    small = copy.copy(cams[0])
    small.width, small.height = 16, 16
    tiny = np.zeros((16, 16), np.float32)
    problems = [(37, [36, 35])]
    opt = hpm.pipeline_options(planar_prior=0, geom_consistency=0, workers=1)
    opt.params.max_image_size = 3200
    states = []
    for filler_cam, filler_img in ((cams[0], imgs[0]), (small, tiny)):
        say(f"case: 38-view pool, fillers {filler_cam.width}x{filler_cam.height}, problem 37 <- 36, 35 at {W}x{H}")
        pool = hpm.Pool([filler_cam] * 35 + list(cams), [filler_img] * 35 + imgs, max_image_size=3200)
        st = hpm.run_pipeline(pool, problems, opt)
        say(f"  {st.runs} RunPatchMatch, {st.n_eval} evaluations")
        states.append(pool.state(37))
        pool.close()
    (h0, c0), (h1, c1) = states
    assert np.array_equal(h0.view(np.uint32), h1.view(np.uint32))
    assert np.array_equal(c0.view(np.uint32), c1.view(np.uint32))
    assert np.mean(c0 < 1.0) > 0.5  # a textured view was matched, not a blank one
