"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bars:
  - shared arithmetic (exp/sin/cos/acos, Philox): bit-identical CPU vs GPU;
  - NCC cost vectors with fixed hypotheses: GPU FP32 within 1e-4 of the
    FP64 evaluation of the reference formula, away from the two decision
    boundaries (var < 1e-5 and the source-image bounds test of the warped
    centre), and bit-identical to the oracle's FP32 arithmetic;
  - each stage of RunPatchMatch (init, 6 half-iterations, finalize, two
    filter passes) started from the same state: bit-identical per pixel;
  - a full RunPatchMatch on the 640x480 four-plane workload: at least 99% of
    pixels within 1% relative depth and 5 degrees normal of the CPU run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_problem(hpm, cams, images, params=None):
    pr = hpm.Problem(0)
    if params is not None:
        pr.set_params(params)
    pr.input_init_memory(cams, images, ref_image_id=0)
    pr.CudaSpaceInitialization()
    return pr


def test_shared_math_bitwise(hpm, cpu):
    say("case: shared math CPU vs GPU")
    import ctypes as C

    L = cpu.lib()
    L.hpm_oracle_fmath.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]
    L.hpm_oracle_philox_block.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    rng = np.random.default_rng(0)
    pr = hpm.Problem(0)
    ranges = {0: (-100.0, 90.0), 1: (-7.0, 7.0), 2: (-7.0, 7.0), 3: (-1.0, 1.0)}
    for kind, (lo, hi) in ranges.items():
        x = rng.uniform(lo, hi, 100000).astype(np.float32)
        x[:5] = np.array([0.0, lo, hi, -0.0, (lo + hi) / 2], np.float32)
        ref = np.empty_like(x)
        L.hpm_oracle_fmath(kind, x.ctypes.data_as(C.POINTER(C.c_float)), ref.ctypes.data_as(C.POINTER(C.c_float)),
                           len(x))
        got = pr.device_selftest(kind, x)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), kind
    # the tap sampling's fraction (5) and floor (6): random values, values
    # just below and at integers, tiny and subnormal negatives, signed zeros
    edge = [np.nextafter(np.float32(k), np.float32(-np.inf)) for k in range(-9, 10)] + list(range(-9, 10))
    edge += [-1e-9, -1e-30, -1e-40, -0.0, 0.0, 1e-40, 0.99999994, -0.99999994, 4194303.5, -4194303.5]
    x = np.concatenate([np.array(edge, np.float32), rng.uniform(-12.0, 12.0, 200000).astype(np.float32),
                        (rng.integers(-2**20, 2**20, 50000) / np.float32(4096.0) - 1e-7).astype(np.float32)])
    for kind in (5, 6):
        ref = np.empty_like(x)
        L.hpm_oracle_fmath(kind, x.ctypes.data_as(C.POINTER(C.c_float)), ref.ctypes.data_as(C.POINTER(C.c_float)),
                           len(x))
        got = pr.device_selftest(kind, x)
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, (kind, x[bad[:5]], got[bad[:5]], ref[bad[:5]])
    # the NCC finish's square root and quotient and the warp's double
    # reciprocal, over their operand ranges and beyond: IEEE results (numpy
    # rounds correctly), bit for bit
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-10), np.log(4.3e9), 300000)),
                        rng.uniform(1e-10, 1e-9, 20000), rng.uniform(1.0, 4.0, 20000)]).astype(np.float32)
    x = np.concatenate([x, np.array([1e-10, 4.3e9, 1.0, 2.0, 0.25, 65025.0**2], np.float32)])
    got = pr.device_selftest(7, x)
    assert np.array_equal(got.view(np.uint32), np.sqrt(x).view(np.uint32))
    num = np.concatenate([rng.uniform(-70000.0, 70000.0, 200000), rng.uniform(-1e-3, 1e-3, 50000),
                          np.exp(rng.uniform(np.log(6e-8), np.log(7e4), 100000)) * rng.choice([-1, 1], 100000),
                          np.zeros(100)]).astype(np.float32)
    den = np.exp(rng.uniform(np.log(1e-5), np.log(65025.0), len(num))).astype(np.float32)
    got = pr.device_selftest(8, np.stack([num, den], 1))
    assert np.array_equal(got, num / den)
    d = np.concatenate([np.exp(rng.uniform(np.log(1e-30), np.log(1e30), 300000)) * rng.choice([-1, 1], 300000),
                        rng.uniform(0.5, 2.0, 50000), np.array([1.0, -1.0, 3.0, 1e-300, 1e300])])
    got = pr.device_selftest(9, d)
    assert np.array_equal(got.view(np.uint64), (1.0 / d).view(np.uint64))
    blocks = rng.integers(0, 2**32, size=(4096, 6), dtype=np.uint64).astype(np.uint32)
    got = pr.device_selftest(4, blocks)
    for i in range(0, len(blocks), 97):
        ctr = (C.c_uint32 * 4)(*[int(v) for v in blocks[i, :4]])
        key = (C.c_uint32 * 2)(*[int(v) for v in blocks[i, 4:]])
        out = (C.c_uint32 * 4)()
        L.hpm_oracle_philox_block(ctr, key, out)
        assert list(out) == [int(v) for v in got[i]]
    pr.close()


def _planes_near_truth(hpm, cams, gd, gn, order, rng, n, noise):
    """Camera-frame planes at random pixels from the ground truth, perturbed."""
    ref = cams[0]
    W, H = ref.width, ref.height
    xs = rng.integers(0, W, n)
    ys = rng.integers(0, H, n)
    R = np.array(ref.R, np.float64).reshape(3, 3)
    K = np.array(ref.K, np.float64).reshape(3, 3)
    d = gd[order[0]][ys, xs].astype(np.float64) * (1.0 + noise * rng.standard_normal(n))
    nw = gn[order[0]][ys, xs].astype(np.float64)
    nc = nw @ R.T
    nc += noise * rng.standard_normal(nc.shape)
    nc /= np.linalg.norm(nc, axis=1, keepdims=True)
    X = np.stack([d * (xs - K[0, 2]) / K[0, 0], d * (ys - K[1, 2]) / K[1, 1], d], axis=1)
    dist = -(nc * X).sum(axis=1)
    planes = np.concatenate([nc, dist[:, None]], axis=1).astype(np.float32)
    pix = np.stack([xs, ys], axis=1).astype(np.int32)
    return planes, pix


def test_cost_vectors_vs_fp64(hpm, cpu, c1_full):
    cfg, (cams, images, gd, gn, low, order) = c1_full
    say("case: cost vectors 640x480, 4 sources, fixed hypotheses")
    rng = np.random.default_rng(1)
    p1, x1 = _planes_near_truth(hpm, cams, gd, gn, order, rng, 20000, 0.01)
    p2, x2 = _planes_near_truth(hpm, cams, gd, gn, order, rng, 20000, 0.3)
    planes = np.concatenate([p1, p2])
    pix = np.concatenate([x1, x2])
    pr = make_problem(hpm, cams, images)
    got = pr.cost_vectors(planes, pix).astype(np.float64)
    params = pr.params()
    o = cpu.Oracle(cams, images, params)
    f32 = o.cost_vectors(planes, pix, fp64=False)
    f64 = o.cost_vectors(planes, pix, fp64=True)
    # bit-identical to the oracle's FP32 arithmetic
    assert np.array_equal(got.astype(np.float32).view(np.uint32), f32.astype(np.float32).view(np.uint32))
    # within 1e-4 of FP64 except where FP32 and FP64 land on different sides
    # of a decision boundary (cost exactly 2 on one side only)
    boundary = (got == 2.0) != (f64 == 2.0)
    err = np.abs(got - f64)[~boundary]
    say(f"max |fp32-fp64| = {err.max():.3e}, boundary flips = {boundary.sum()} of {boundary.size}")
    assert err.max() <= 1e-4
    assert boundary.mean() < 1e-3
    pr.close()


def test_cost_vectors_fractional_images_and_borders(hpm, cpu, c1_full):
    """Source texels are sampled as their nearest FP16 value (exact for 8-bit
    grey images). Images with fractional grey values, which FP16 rounds, and
    hypotheses at every pixel within 8 of the image border, where windows
    leave the source image and the clamp decides: bit-identical to the
    oracle."""
    cfg, (cams, images, gd, gn, low, order) = c1_full
    say("case: cost vectors 640x480, fractional images, border pixels")
    frac = [(im * np.float32(0.731) + np.float32(0.1234)).astype(np.float32) for im in images]
    assert any(not np.array_equal(f.astype(np.float16).astype(np.float32), f) for f in frac)
    W, H = cams[0].width, cams[0].height
    yy, xx = np.mgrid[0:H, 0:W]
    border = (xx < 8) | (xx >= W - 8) | (yy < 8) | (yy >= H - 8)
    rng = np.random.default_rng(4)
    p1, x1 = _planes_near_truth(hpm, cams, gd, gn, order, rng, 20000, 0.05)
    bx, by = xx[border], yy[border]
    R = np.array(cams[0].R, np.float64).reshape(3, 3)
    K = np.array(cams[0].K, np.float64).reshape(3, 3)
    d = gd[order[0]][by, bx].astype(np.float64) * (1.0 + 0.05 * rng.standard_normal(bx.size))
    nc = gn[order[0]][by, bx].astype(np.float64) @ R.T + 0.05 * rng.standard_normal((bx.size, 3))
    nc /= np.linalg.norm(nc, axis=1, keepdims=True)
    X = np.stack([d * (bx - K[0, 2]) / K[0, 0], d * (by - K[1, 2]) / K[1, 1], d], axis=1)
    p2 = np.concatenate([nc, -(nc * X).sum(axis=1)[:, None]], axis=1).astype(np.float32)
    x2 = np.stack([bx, by], axis=1).astype(np.int32)
    planes = np.concatenate([p1, p2])
    pix = np.concatenate([x1, x2])
    pr = make_problem(hpm, cams, frac)
    got = pr.cost_vectors(planes, pix)
    o = cpu.Oracle(cams, frac, pr.params())
    want = o.cost_vectors(planes, pix, fp64=False)
    same = got.astype(np.float32).view(np.uint32) == want.astype(np.float32).view(np.uint32)
    say(f"identical {same.mean():.6f} of {same.size}; valid costs {(want < 2).mean():.3f}")
    assert same.all()
    pr.close()


def test_stages_bitwise(hpm, cpu, c1_small):
    """Every stage started from the oracle's state gives the oracle's bits."""
    cfg, (cams, images, gd, gn, low, order) = c1_small
    say("case: stagewise 160x120, 4 sources")
    pr = make_problem(hpm, cams, images)
    o = cpu.Oracle(cams, images, pr.params(), salt=0)

    def compare(tag):
        hyp, cost, sel = pr.debug_state()
        same = (np.all(hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
                & ((cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(cost) & np.isnan(o.cost)))
                & (sel == o.sel))
        say(f"{tag}: identical pixels {same.mean():.6f}")
        assert same.all(), tag

    def sync():
        import ctypes as C
        hpm.lib().hpm_debug_write_state(pr._h, hpm.fptr(o.hyp), hpm.fptr(o.cost), hpm.uptr(o.sel))

    o.init()
    pr.debug_stage(0)
    compare("init")
    for it in range(3):
        for colour in range(2):
            sync()
            o.half_iteration(it, colour)
            pr.debug_stage(1, it, colour)
            compare(f"iter {it} colour {colour}")
    sync()
    o.finalize()
    pr.debug_stage(2)
    compare("finalize")
    for colour in range(2):
        sync()
        o.filter(colour)
        pr.debug_stage(3, 0, colour)
        compare(f"filter {colour}")
    pr.close()


def test_state_write_over_initialised_problem(hpm, cpu, c1_small):
    """The device keeps each pixel's per-view costs of its hypothesis for the
    rescore; a state written over an initialised problem (other planes,
    costs and views) must propagate as the oracle does from that state, over
    several half-iterations without a further write."""
    cfg, (cams, images, gd, gn, low, order) = c1_small
    say("case: state write over an initialised problem, 160x120, 4 sources")
    pr = make_problem(hpm, cams, images)
    pr.debug_stage(0)
    dev_hyp, _, _ = pr.debug_state()
    o = cpu.Oracle(cams, images, pr.params(), salt=0)
    other = cpu.Oracle(cams, images, pr.params(), salt=1)
    other.init()
    o.hyp[...] = other.hyp
    o.cost[...] = other.cost
    o.sel[...] = other.sel
    moved = np.any(o.hyp.view(np.uint32) != dev_hyp.view(np.uint32), axis=2).mean()
    say(f"written planes differ from the device's at {moved:.4f} of the pixels")
    assert moved > 0.9
    hpm.lib().hpm_debug_write_state(pr._h, hpm.fptr(o.hyp), hpm.fptr(o.cost), hpm.uptr(o.sel))
    for it, colour in ((0, 0), (0, 1), (1, 0)):
        o.half_iteration(it, colour)
        pr.debug_stage(1, it, colour)
        hyp, cost, sel = pr.debug_state()
        same = (np.all(hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
                & ((cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(cost) & np.isnan(o.cost)))
                & (sel == o.sel))
        say(f"iter {it} colour {colour}: identical pixels {same.mean():.6f}")
        assert same.all()
    pr.close()


def depth_normal_agreement(a_hyp, b_hyp):
    da, db = a_hyp[..., 3], b_hyp[..., 3]
    rel = np.abs(da - db) / np.maximum(np.abs(db), 1e-12)
    na, nb = a_hyp[..., :3], b_hyp[..., :3]
    cosang = np.clip((na * nb).sum(-1) / np.maximum(np.linalg.norm(na, axis=-1) * np.linalg.norm(nb, axis=-1),
                                                    1e-12), -1, 1)
    ang = np.degrees(np.arccos(cosang))
    ok = (rel < 0.01) & (ang < 5.0)
    return ok


def test_full_run_parity_c1(hpm, cpu, c1_full):
    cfg, (cams, images, gd, gn, low, order) = c1_full
    say("case: full RunPatchMatch 640x480, 4 sources")
    pr = make_problem(hpm, cams, images)
    pr.RunPatchMatch()
    g_hyp = pr.hypotheses()
    g_cost = pr.costs()
    stats = pr.run_stats()
    o = cpu.Oracle(cams, images, pr.params(), salt=0)
    o.run_patchmatch()
    ok = depth_normal_agreement(g_hyp, o.hyp)
    ident = np.all(g_hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2).mean()
    gt = gd[order[0]]
    acc_g = (np.abs(g_hyp[..., 3] - gt) / gt < 0.02).mean()
    acc_o = (np.abs(o.hyp[..., 3] - gt) / gt < 0.02).mean()
    say(f"agree(1%,5deg)={ok.mean():.5f} identical={ident:.5f} acc2%: gpu={acc_g:.4f} cpu={acc_o:.4f} "
        f"n_eval gpu={stats.n_eval} cpu={o.n_eval} ms={stats.ms_total:.2f}")
    same_c = (g_cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(g_cost) & np.isnan(o.cost))
    assert ok.mean() >= 0.99
    assert ident == 1.0 and same_c.all()
    assert stats.n_eval == o.n_eval
    assert acc_g > 0.8
    pr.close()


@pytest.mark.parametrize("top_k", [1, 8, 9, 12])
def test_init_top_k_bitwise(hpm, cpu, top_k):
    """The initial cost (mean of the top_k smallest view costs) and view mask
    with 12 sources: top_k up to 8 keeps the smallest costs in registers,
    above 8 sorts them; both give the oracle's bits, and one half-iteration
    from that state too."""
    from conftest import problem_inputs

    w, h, nv = 64, 48, 13
    f = 0.8 * w
    cfg = hpm.scene_config(seed=71, width=w, height=h, fx=f, fy=f, cx=w / 2.0, cy=h / 2.0, num_views=nv)
    cams, images, gd, gn, low, order = problem_inputs(hpm, cfg, 6)
    say(f"case: init with top_k {top_k}: {w}x{h}, {len(cams) - 1} sources")
    p = hpm.default_params()
    p.top_k = top_k
    pr = make_problem(hpm, cams, images, p)
    o = cpu.Oracle(cams, images, pr.params(), salt=0)
    o.init()
    pr.debug_stage(0)
    for tag in ("init", "iter 0 colour 0"):
        if tag != "init":
            o.half_iteration(0, 0)
            pr.debug_stage(1, 0, 0)
        hyp, cost, sel = pr.debug_state()
        same = (np.all(hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
                & ((cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(cost) & np.isnan(o.cost)))
                & (sel == o.sel))
        say(f"{tag}: identical pixels {same.mean():.6f}, mean views selected {np.mean([bin(int(s)).count('1') for s in sel.ravel()]):.2f}")
        assert same.all(), tag
    pr.close()


def _resized_view(hpm, cam, img, w, h):
    """a source view at another resolution: the image resized, K scaled"""
    import copy

    c = copy.copy(cam)
    sx, sy = w / cam.width, h / cam.height
    K = list(c.K)
    K[0] *= sx
    K[2] *= sx
    K[4] *= sy
    K[5] *= sy
    for i in range(9):
        c.K[i] = K[i]
    c.width, c.height = w, h
    return c, hpm.resize_bilinear(img, w, h)


@pytest.mark.parametrize("case", ["one_source", "max_sources", "tiny", "odd", "mixed_sizes"])
def test_full_run_edge_cases(hpm, cpu, case):
    """Full RunPatchMatch against the oracle, bit for bit (hypotheses, costs,
    evaluation count), on shapes at the edges of the problem: a single
    source, the maximum of 32 sources, an image smaller than the 11x11
    window, odd sizes, and sources at other resolutions than the reference."""
    from conftest import problem_inputs

    sizes = {"one_source": (40, 30, 2), "max_sources": (48, 36, 33), "tiny": (9, 7, 5), "odd": (97, 53, 5),
             "mixed_sizes": (96, 72, 5)}
    w, h, nv = sizes[case]
    f = 0.8 * w
    cfg = hpm.scene_config(seed=40 + len(case), width=w, height=h, fx=f, fy=f, cx=w / 2.0, cy=h / 2.0, num_views=nv)
    cams, images, gd, gn, low, order = problem_inputs(hpm, cfg, nv // 2)
    if case == "mixed_sizes":
        cams[1], images[1] = _resized_view(hpm, cams[1], images[1], 64, 48)
        cams[3], images[3] = _resized_view(hpm, cams[3], images[3], 120, 90)
    say(f"case: full run {case}: {w}x{h}, {len(cams) - 1} sources")
    pr = make_problem(hpm, cams, images)
    pr.RunPatchMatch()
    g_hyp, g_cost = pr.hypotheses(), pr.costs()
    stats = pr.run_stats()
    o = cpu.Oracle(cams, images, pr.params(), salt=0)
    o.run_patchmatch()
    same_h = np.all(g_hyp.view(np.uint32) == o.hyp.view(np.uint32), axis=2)
    same_c = (g_cost.view(np.uint32) == o.cost.view(np.uint32)) | (np.isnan(g_cost) & np.isnan(o.cost))
    say(f"identical hypotheses {same_h.mean():.5f}, costs {same_c.mean():.5f}, n_eval gpu={stats.n_eval} "
        f"cpu={o.n_eval}")
    assert same_h.all() and same_c.all()
    assert stats.n_eval == o.n_eval
    pr.close()
