"""Folder-format I/O (cams, pair.txt, .dmb, PGM), the bilinear resize, and
the synthetic scene generator's ground truth."""
import os

import numpy as np
import pytest


def test_dmb_round_trip_and_header(hpm, tmp_path):
    rng = np.random.default_rng(0)
    d = rng.normal(size=(7, 5)).astype(np.float32)
    n = rng.normal(size=(7, 5, 3)).astype(np.float32)
    hpm.write_dmb(str(tmp_path / "d.dmb"), d)
    hpm.write_dmb(str(tmp_path / "n.dmb"), n)
    assert np.array_equal(hpm.read_dmb(str(tmp_path / "d.dmb")), d)
    assert np.array_equal(hpm.read_dmb(str(tmp_path / "n.dmb")), n)
    raw = open(tmp_path / "n.dmb", "rb").read()
    assert np.frombuffer(raw[:16], np.int32).tolist() == [1, 7, 5, 3]  # type, h, w, nb (HPM.cpp:376-400)
    assert len(raw) == 16 + 7 * 5 * 3 * 4
    # type != 1 is rejected (HPM.cpp:315-318)
    bad = bytearray(raw)
    bad[0] = 2
    open(tmp_path / "bad.dmb", "wb").write(bytes(bad))
    with pytest.raises(hpm.HpmError):
        hpm.read_dmb(str(tmp_path / "bad.dmb"))
    with pytest.raises(hpm.HpmError):
        hpm.read_dmb(str(tmp_path / "missing.dmb"))


def test_camera_file(hpm, tmp_path):
    text = """extrinsic
0.9 0.1 0.0 1.5
-0.1 0.9 0.2 -2
0 0 1 3
0 0 0 1

intrinsic
500 0 320
0 510 240
0 0 1

2.5 0.01 192 8.5
"""
    p = tmp_path / "00000003_cam.txt"
    p.write_text(text)
    c = hpm.read_camera(str(p))
    assert list(c.R) == pytest.approx([0.9, 0.1, 0.0, -0.1, 0.9, 0.2, 0, 0, 1])
    assert list(c.t) == pytest.approx([1.5, -2, 3])
    assert list(c.K) == pytest.approx([500, 0, 320, 0, 510, 240, 0, 0, 1])
    assert (c.depth_min, c.depth_max) == pytest.approx((2.5, 8.5))
    hpm.write_camera(str(tmp_path / "w.txt"), c)
    c2 = hpm.read_camera(str(tmp_path / "w.txt"))
    for f in ("K", "R", "t"):
        assert list(getattr(c2, f)) == list(getattr(c, f))
    assert (c2.depth_min, c2.depth_max) == (c.depth_min, c.depth_max)


def test_pair_file_skips_nonpositive_scores(hpm, tmp_path):
    p = tmp_path / "pair.txt"
    p.write_text("3\n0\n3 1 10.0 2 0.0 5 -1\n1\n2 0 3.5 2 1.0\n2\n0\n")
    probs = hpm.read_pair(str(p))
    assert probs == [(0, [1]), (1, [0, 2]), (2, [])]  # main.cpp:26-28


def test_pgm_round_trip(hpm, tmp_path):
    img = (np.arange(35 * 17) % 256).astype(np.uint8).reshape(17, 35)
    path = str(tmp_path / "a.pgm")
    assert hpm.lib().hpm_write_pgm_u8(path.encode(), hpm.bptr(img), 35, 17) == 0
    assert np.array_equal(hpm.read_pgm(path), img.astype(np.float32))


def test_resize_bilinear_half_pixel(hpm):
    # constant stays constant; a horizontal ramp halves exactly at half-pixel centres
    c = np.full((10, 12), 3.5, np.float32)
    assert np.all(hpm.resize_bilinear(c, 6, 5) == 3.5)
    ramp = np.tile(np.arange(12, dtype=np.float32), (4, 1))
    out = hpm.resize_bilinear(ramp, 6, 4)
    # dst x -> src (x + 0.5) * 2 - 0.5 = 2x + 0.5
    assert np.allclose(out[0], 2 * np.arange(6) + 0.5)
    up = hpm.resize_bilinear(ramp, 24, 4)
    src = (np.arange(24) + 0.5) * 0.5 - 0.5
    assert np.allclose(up[0], np.clip(src, 0, 11))


def test_scene_is_deterministic_and_consistent(hpm):
    cfg = hpm.scene_config(seed=5, width=200, height=150, fx=160.0, fy=160.0, cx=100.0, cy=75.0, num_views=3,
                           num_planes=3, num_boxes=4, low_texture_fraction=0.3)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    cams2, imgs2, gd2, gn2, low2 = hpm.render_scene(cfg)
    assert np.array_equal(imgs, imgs2) and np.array_equal(gd, gd2)
    assert 0.15 < low.mean() < 0.6
    # every middle-view pixel hits geometry inside the depth slab
    assert gd[1].min() >= cfg.depth_min - 1e-4 and gd[1].max() <= cfg.depth_max + 1e-4
    # normals are unit and face the camera
    nrm = gn[1].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    # cross-view consistency: project view 1 points into view 2 and compare depth
    def KRt(c):
        return (np.array(c.K, np.float64).reshape(3, 3), np.array(c.R, np.float64).reshape(3, 3),
                np.array(c.t, np.float64))
    K1, R1, t1 = KRt(cams[1])
    K2, R2, t2 = KRt(cams[2])
    H, W = gd[1].shape
    ys, xs = np.mgrid[0:H, 0:W]
    d = gd[1].astype(np.float64)
    Xc = np.stack([(xs - K1[0, 2]) / K1[0, 0] * d, (ys - K1[1, 2]) / K1[1, 1] * d, d], -1)
    Xw = (Xc - t1) @ R1  # R^T (x - t)
    X2 = Xw @ R2.T + t2
    u = K2[0, 0] * X2[..., 0] / X2[..., 2] + K2[0, 2]
    v = K2[1, 1] * X2[..., 1] / X2[..., 2] + K2[1, 2]
    ui, vi = np.rint(u).astype(int), np.rint(v).astype(int)
    ok = (ui >= 0) & (ui < W) & (vi >= 0) & (vi < H) & (np.abs(u - ui) < 0.05) & (np.abs(v - vi) < 0.05)
    d2 = gd[2][vi[ok], ui[ok]]
    rel = np.abs(d2 - X2[..., 2][ok]) / d2
    assert np.mean(rel < 1e-3) > 0.8  # the rest are occlusions and silhouettes
    # normal convention: world normal, facing the camera centre
    C1 = -R1.T @ t1
    pts = Xw.reshape(-1, 3)
    facing = ((C1 - pts) * nrm).sum(1)
    assert np.mean(facing > 0) > 0.999


def test_scene_folder_layout(hpm, tmp_path):
    cfg = hpm.scene_config(seed=3, width=64, height=48, fx=50.0, fy=50.0, cx=32.0, cy=24.0, num_views=3)
    hpm.write_scene_folder(cfg, str(tmp_path))
    for i in range(3):
        assert os.path.exists(tmp_path / "images" / f"{i:08d}.pgm")
        assert os.path.exists(tmp_path / "cams" / f"{i:08d}_cam.txt")
        assert hpm.read_dmb(str(tmp_path / "gt" / f"{i:08d}_depth.dmb")).shape == (48, 64)
    probs = hpm.read_pair(str(tmp_path / "pair.txt"))
    assert probs == [(0, [1, 2]), (1, [0, 2]), (2, [1, 0])]
