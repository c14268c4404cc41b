"""The batched form (hpm_batch.h) against unbatched RunPatchMatch calls.

Bar: every problem of a batch ends with hypotheses and costs bit-identical
to hpm_run_patchmatch on its own handle, on the first call (autotuned and
captured) and on a replay of the captured schedule (the second RunPatchMatch
call of each problem, with the next RNG stream), for problems of different
sizes and source counts, in plain and geometric-consistency mode.
"""
import numpy as np
import pytest

from conftest import problem_inputs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _scenes(hpm):
    shapes = [(160, 120, 5, 11), (224, 136, 4, 12), (96, 72, 6, 13)]
    out = []
    for w, h, n, seed in shapes:
        f = 0.8 * w
        cfg = hpm.scene_config(seed=seed, width=w, height=h, fx=f, fy=f, cx=w / 2.0, cy=h / 2.0, num_views=n)
        cams, imgs, gd, gn, low, order = problem_inputs(hpm, cfg, n // 2)
        out.append((cams, imgs, gd, gn, order))
    return out


def _problem(hpm, cams, imgs, depths=None):
    pr = hpm.Problem(0)
    if depths is not None:
        pr.SetGeomConsistencyParams(False)
    pr.input_init_memory(cams, imgs, depths=depths, ref_image_id=0)
    pr.CudaSpaceInitialization()
    return pr


@pytest.mark.parametrize("geom", [False, True])
def test_batch_matches_unbatched(hpm, geom):
    say(f"case: batch of 3 problems (160x120x5, 224x136x4, 96x72x6), geometric={geom}")
    scenes = _scenes(hpm)
    singles, batched = [], []
    for cams, imgs, gd, gn, order in scenes:
        depths = [gd[i] * np.float32(1.01) for i in order] if geom else None
        a = _problem(hpm, cams, imgs, depths)
        b = _problem(hpm, cams, imgs, depths)
        if geom:
            # a geometric run starts from a previous state: ground truth here
            ref = order[0]
            hyp = np.concatenate([gn[ref], gd[ref][..., None]], axis=-1).astype(np.float32)
            cost = np.full(gd[ref].shape, 0.5, np.float32)
            a.device_init_state(hyp, cost)
            b.device_init_state(hyp, cost)
        singles.append(a)
        batched.append(b)
    batch = hpm.Batch(0)
    for call in range(2):
        for pr in singles:
            pr.RunPatchMatch()
        st = batch.run(batched)
        say(f"call {call}: grid {st.grid}, tiles {st.tiles_full}/{st.tiles_colour}, reused {st.graph_reused}, "
            f"{st.ms_device:.2f} ms device, {st.n_eval} evals")
        assert st.graph_reused == (1 if call else 0)
        for i, (a, b) in enumerate(zip(singles, batched)):
            ha, hb = a.hypotheses(), b.hypotheses()
            ca, cb = a.costs(), b.costs()
            assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32)), (call, i)
            assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32)), (call, i)
            assert a.run_stats().n_eval == b.run_stats().n_eval, (call, i)
    log = batch.log()
    say(log.strip())
    assert "chosen" in log and "queue items of 2 tiles" in log
    # every schedule the tuner may choose gives the same bits
    for items, persistent in ((0, False), (1, False), (2, False), (4, False), (1, True), (2, True)):
        batch.set_schedule(0, items, persistent)
        for pr in singles:
            pr.RunPatchMatch()
        st = batch.run(batched)
        say(f"fixed schedule, items {items}, persistent {persistent}: grid {st.grid}, {st.ms_device:.2f} ms device")
        assert st.items == items and st.persistent == int(persistent)
        for i, (a, b) in enumerate(zip(singles, batched)):
            assert np.array_equal(a.hypotheses().view(np.uint32), b.hypotheses().view(np.uint32)), (items, i)
            assert np.array_equal(a.costs().view(np.uint32), b.costs().view(np.uint32)), (items, i)
    batch.close()
    for pr in singles + batched:
        pr.close()
