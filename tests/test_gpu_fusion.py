"""Fusion on the device (hpm_fusion_run: the consistency tests of every
(pixel, source) pair in hpm_k_fusion_geometry, the masks replayed on the
host in the reference's order) against the oracle's sequential RunFusion
(main.cpp:502-633) on the same maps. Bar: the same points in the same
order, bit for bit."""
import numpy as np
import pytest

from test_gpu_parity import say
from test_gpu_pipeline import scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geom", [True, False])
def test_fusion_matches_oracle(hpm, cpu, geom, tmp_path):
    cfg, cams, imgs, gd, problems = scene(hpm)
    say(f"case: fusion of the 4-view 320x240 scene, geometric maps {geom}")
    pool = hpm.Pool(cams, imgs)
    opt = hpm.pipeline_options()
    opt.params.geom_iterations = 1
    opt.geom_consistency = 1 if geom else 0
    hpm.run_pipeline(pool, problems, opt)
    probs = [(v, list(s)) for v, s in problems]
    got = hpm.fusion(pool, probs, geom_consistency=geom)
    depths = [pool.depth(v, 1 if geom else 0) for v, _ in probs]
    normals = [pool.state(v)[0][..., :3].copy() for v, _ in probs]
    colors = [np.repeat(np.clip(im, 0, 255).astype(np.uint8)[..., None], 3, axis=2) for im in imgs]
    cam_list = [pool.camera(v) for v, _ in probs]
    want = cpu.fusion(cam_list, depths, normals, colors, probs, hpm.POINT_DTYPE)
    say(f"points: device {len(got)}, oracle {len(want)} of {sum(d.size for d in depths)} pixels")
    assert len(got) == len(want) and len(got) > 0
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    # the kept points lie on the scene: depth error against ground truth
    path = str(tmp_path / "fused.ply")
    hpm.write_ply(path, got)
    assert (tmp_path / "fused.ply").stat().st_size > 27 * len(got)
    pool.close()
