"""COLMAP sparse model -> images/, cams/, pair.txt (hpm-mvs_amd/colmap_to_mvs.py,
the conversion of ACMM+NESP/windows/colmap2mvsnet_acm.py:280-451).

A synthetic model is built from known cameras and points: world-to-camera
rotations from a look-at construction, written as quaternions by scipy (so
the converter's quaternion convention is checked against an independent
routine, with random quaternion signs), four camera models, 2D points with
and without a 3D point, and one image 1 mm from another (their pair must
score 0: triangulation angles far below 1 degree). The converter's output is
read back with the pipeline's own readers and compared with the ground
truth: poses and K, the depth ranges (percentiles of the observed points'
depths), the 1-pixel inverse-depth count against its closed form
(depth_min / f of parallax at depth_min), the pair scores against a brute
force over the tracks, and the images (.jpg copied byte for byte, .png
re-encoded). The text and binary model formats give identical files.
"""
import os
import struct

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import load_pkg

W, H = 320, 240
N = 7
CAMERAS = {  # id: (model, params)
    1: ("SIMPLE_PINHOLE", [300.0, 160.0, 120.0]),
    2: ("PINHOLE", [310.0, 305.0, 158.5, 121.25]),
    3: ("OPENCV", [290.0, 295.0, 161.0, 119.0, 0.01, -0.02, 0.001, 0.0005]),
    4: ("SIMPLE_RADIAL", [305.0, 159.0, 118.0, 0.03]),
}
MODEL_ID = {"SIMPLE_PINHOLE": 0, "PINHOLE": 1, "SIMPLE_RADIAL": 2, "OPENCV": 4}


def K_of(cid):
    model, p = CAMERAS[cid]
    if model in ("SIMPLE_PINHOLE", "SIMPLE_RADIAL"):
        return np.array([[p[0], 0, p[1]], [0, p[0], p[2]], [0, 0, 1]])
    return np.array([[p[0], 0, p[2]], [0, p[1], p[3]], [0, 0, 1]])


def look_at(c, target, roll):
    z = (target - c) / np.linalg.norm(target - c)
    x = np.cross([0.0, 1.0, 0.0], z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    cr, sr = np.cos(roll), np.sin(roll)
    return np.array([[cr, -sr, 0], [sr, cr, 0], [0, 0, 1]]) @ R


def make_model(rng):
    pts = np.stack([rng.uniform(-1, 1, 300), rng.uniform(-0.7, 0.7, 300), rng.uniform(4, 6, 300)], 1)
    centers = [np.array([0.3 * i - 0.9, 0.05 * i, 0.0]) for i in range(N - 1)]
    centers.append(centers[0] + np.array([0.001, 0.0, 0.0]))  # 1 mm from image 1
    images = []
    for i, c in enumerate(centers):
        R = look_at(c, np.array([0.0, 0.0, 5.0]), rng.uniform(-0.1, 0.1))
        t = -R @ c
        cid = i % 4 + 1
        K = K_of(cid)
        obs = []  # (x, y, point id)
        for pid, X in enumerate(pts):
            p = K @ (R @ X + t)
            u, v = p[0] / p[2], p[1] / p[2]
            if p[2] > 0 and 0 <= u < W and 0 <= v < H and rng.random() < 0.7:
                obs.append((u, v, pid + 1))
        for _ in range(15):
            obs.append((rng.uniform(0, W), rng.uniform(0, H), -1))
        rng.shuffle(obs)
        xyzw = Rotation.from_matrix(R).as_quat()
        q = np.array([xyzw[3], xyzw[0], xyzw[1], xyzw[2]]) * (1 if rng.random() < 0.5 else -1)
        name = f"view_{i}.png" if i == 2 else f"view_{i}.jpg"
        # image ids out of order: the converter numbers images by sorted id
        images.append(dict(id=10 * (N - i), q=q, t=t, R=R, cid=cid, name=name, obs=obs))
    return pts, images


def write_text(sparse, pts, images):
    with open(os.path.join(sparse, "cameras.txt"), "w") as f:
        f.write("# Camera list\n")
        for cid, (model, p) in CAMERAS.items():
            f.write(f"{cid} {model} {W} {H} " + " ".join(repr(v) for v in p) + "\n")
    with open(os.path.join(sparse, "images.txt"), "w") as f:
        f.write("# Image list with two lines of data per image\n")
        for im in images:
            f.write(f"{im['id']} " + " ".join(repr(float(v)) for v in (*im['q'], *im['t'])) +
                    f" {im['cid']} {im['name']}\n")
            f.write(" ".join(f"{float(u)!r} {float(v)!r} {pid}" for u, v, pid in im["obs"]) + "\n")
    with open(os.path.join(sparse, "points3D.txt"), "w") as f:
        f.write("# 3D point list\n")
        for pid, X in enumerate(pts, start=1):
            track = [(im["id"], k) for im in images for k, o in enumerate(im["obs"]) if o[2] == pid]
            f.write(f"{pid} {float(X[0])!r} {float(X[1])!r} {float(X[2])!r} 128 128 128 0.5 " +
                    " ".join(f"{a} {b}" for a, b in track) + "\n")


def write_binary(sparse, pts, images):
    with open(os.path.join(sparse, "cameras.bin"), "wb") as f:
        f.write(struct.pack("<Q", len(CAMERAS)))
        for cid, (model, p) in CAMERAS.items():
            f.write(struct.pack("<iiQQ", cid, MODEL_ID[model], W, H) + struct.pack("<" + "d" * len(p), *p))
    with open(os.path.join(sparse, "images.bin"), "wb") as f:
        f.write(struct.pack("<Q", len(images)))
        for im in images:
            f.write(struct.pack("<i4d3di", im["id"], *im["q"], *im["t"], im["cid"]) + im["name"].encode() + b"\0")
            f.write(struct.pack("<Q", len(im["obs"])))
            for u, v, pid in im["obs"]:
                f.write(struct.pack("<ddq", u, v, pid))
    with open(os.path.join(sparse, "points3D.bin"), "wb") as f:
        f.write(struct.pack("<Q", len(pts)))
        for pid, X in enumerate(pts, start=1):
            track = [(im["id"], k) for im in images for k, o in enumerate(im["obs"]) if o[2] == pid]
            f.write(struct.pack("<Q3d3Bd", pid, *X, 128, 128, 128, 0.5) + struct.pack("<Q", len(track)))
            for a, b in track:
                f.write(struct.pack("<ii", a, b))


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    from PIL import Image as PILImage

    rng = np.random.default_rng(5)
    pts, images = make_model(rng)
    root = tmp_path_factory.mktemp("colmap")
    dense = root / "dense"
    os.makedirs(dense / "sparse")
    os.makedirs(dense / "images")
    write_text(str(dense / "sparse"), pts, images)
    write_binary(str(dense / "sparse"), pts, images)
    yy, xx = np.mgrid[0:H, 0:W]
    for i, im in enumerate(images):
        rgb = np.stack([(xx * (i + 1)) % 256, (yy * 2) % 256, (xx + yy) % 256], -1).astype(np.uint8)
        path = dense / "images" / im["name"]
        PILImage.fromarray(rgb).save(path, **({"quality": 90} if im["name"].endswith(".jpg") else {}))
        im["rgb"] = rgb
    # images in the order the converter numbers them (ascending image id)
    images.sort(key=lambda im: im["id"])
    return dense, root, pts, images


def convert(dense, out, **kw):
    load_pkg()
    from hpm_mvs_amd import colmap_to_mvs

    os.makedirs(out, exist_ok=True)
    return colmap_to_mvs.convert(str(dense), str(out), **kw)


def read_depth_line(path):
    return [float(v) for v in open(path).read().split()[-4:]]


def brute_scores(pts, images):
    centers = [-im["R"].T @ im["t"] for im in images]
    score = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            if i == j:
                continue
            shared = {o[2] for o in images[i]["obs"]} & {o[2] for o in images[j]["obs"]} - {-1}
            ang = []
            for pid in shared:
                a, b = centers[i] - pts[pid - 1], centers[j] - pts[pid - 1]
                ang.append(np.degrees(np.arccos(np.clip(a @ b / np.linalg.norm(a) / np.linalg.norm(b), -1, 1))))
            score[i, j] = 0 if (not ang or np.sort(ang)[int(0.75 * len(ang))] < 1.0) else len(ang)
    return score


@pytest.mark.parametrize("ext", [".txt", ".bin"])
def test_cameras_and_depth_ranges(hpm, scene, ext):
    dense, root, pts, images = scene
    out = root / ("out" + ext)
    assert convert(dense, out, model_ext=ext) == N
    for i, im in enumerate(images):
        path = str(out / "cams" / ("%08d_cam.txt" % i))
        cam = hpm.read_camera(path)
        assert np.allclose(np.array(cam.R).reshape(3, 3), im["R"], atol=1e-6)
        assert np.allclose(np.array(cam.t), im["t"], atol=1e-5)
        assert np.allclose(np.array(cam.K).reshape(3, 3), K_of(im["cid"]), atol=1e-4)
        words = open(path).read().split()
        E = np.array(words[1:17], np.float64).reshape(4, 4)
        assert np.allclose(E[:3, :3], im["R"], atol=1e-12) and np.allclose(E[3], [0, 0, 0, 1])
        z = np.sort([(im["R"] @ pts[pid - 1] + im["t"])[2] for _, _, pid in im["obs"] if pid > 0])
        dmin, interval, num, dmax = read_depth_line(path)
        assert dmin == pytest.approx(z[int(0.01 * len(z))] * 0.75, abs=1e-6)
        assert dmax == pytest.approx(z[int(0.99 * len(z))] * 1.25, abs=1e-6)
        assert num == 192 and interval == pytest.approx((dmax - dmin) / 191, abs=2e-6)
        assert cam.depth_min == pytest.approx(dmin, rel=1e-6) and cam.depth_max == pytest.approx(dmax, rel=1e-6)


def test_inverse_depth_count(scene):
    """max_d 0: the depth count of one pixel of parallax at depth_min, whose
    closed form for a zero-skew K is a step of depth_min / fx"""
    dense, root, pts, images = scene
    out = root / "out_invd"
    convert(dense, out, max_d=0, interval_scale=2.0)
    for i, im in enumerate(images):
        dmin, interval, num, dmax = read_depth_line(str(out / "cams" / ("%08d_cam.txt" % i)))
        fx = K_of(im["cid"])[0, 0]
        want = (1 / dmin - 1 / dmax) / (1 / dmin - 1 / (dmin + dmin / fx))
        # the file holds 6 decimals of each value
        assert num == pytest.approx(want, rel=1e-6)
        assert interval == pytest.approx((dmax - dmin) / (want - 1) / 2.0, rel=1e-4)


def test_pair_scores(hpm, scene):
    dense, root, pts, images = scene
    out = root / "out.txt"
    convert(dense, out)
    score = brute_scores(pts, images)
    # the 1 mm pair shares many points but scores 0
    near = [k for k, im in enumerate(images) if im["name"] in ("view_0.jpg", "view_6.jpg")]
    shared = {o[2] for o in images[near[0]]["obs"]} & {o[2] for o in images[near[1]]["obs"]}
    assert len(shared) > 50 and score[near[0], near[1]] == 0
    words = open(out / "pair.txt").read().split()
    assert int(words[0]) == N
    k = 1
    listed = []
    for i in range(N):
        assert int(words[k]) == i and int(words[k + 1]) == N - 1
        pairs = [(int(words[k + 2 + 2 * m]), int(words[k + 3 + 2 * m])) for m in range(N - 1)]
        k += 2 + 2 * (N - 1)
        # the ranking runs over the whole score row, the image's own 0
        # included, so an image with a zero-score partner may list itself
        assert len({j for j, _ in pairs}) == N - 1
        assert [s for _, s in pairs] == sorted(score[i].astype(int), reverse=True)[:N - 1]
        assert all(s == score[i, j] for j, s in pairs)
        if i in near:
            assert i in [j for j, _ in pairs] or (near[0] + near[1] - i) in [j for j, _ in pairs]
        listed.append([j for j, s in pairs if s > 0])
    # the pipeline's reader keeps the sources of positive score
    parsed = hpm.read_pair(str(out / "pair.txt"))
    assert [r for r, _ in parsed] == list(range(N))
    assert [s for _, s in parsed] == listed


def test_text_and_binary_models_agree(scene):
    dense, root, pts, images = scene
    a, b = root / "agree_txt", root / "agree_bin"
    convert(dense, a, model_ext=".txt")
    convert(dense, b, model_ext=".bin")
    files = ["pair.txt"] + [os.path.join("cams", "%08d_cam.txt" % i) for i in range(N)]
    for name in files:
        assert open(a / name).read() == open(b / name).read(), name


def test_images_copied_or_reencoded(hpm, scene):
    dense, root, pts, images = scene
    out = root / "out_img"
    convert(dense, out)
    for i, im in enumerate(images):
        dst = out / "images" / ("%08d.jpg" % i)
        if im["name"].endswith(".jpg"):
            assert open(dst, "rb").read() == open(dense / "images" / im["name"], "rb").read()
        else:
            rgb = hpm.read_image_rgb(str(dst))
            assert rgb.shape == (H, W, 3)
            assert np.abs(rgb.astype(int) - im["rgb"].astype(int)).mean() < 8.0
