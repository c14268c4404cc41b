"""Converts a COLMAP sparse reconstruction into the folder layout the
pipeline reads (images/%08d.jpg, cams/%08d_cam.txt, pair.txt): the
conversion of ACMM+NESP/windows/colmap2mvsnet_acm.py:280-451, written
without OpenCV.

    python hpm-mvs_amd/colmap_to_mvs.py --dense_folder D --save_folder S [--model_ext .bin]

D holds images/ and sparse/ (cameras, images, points3D as .txt or .bin,
COLMAP's formats). For every image, in image-id order, S receives:
  - cams/%08d_cam.txt: the world-to-camera [R | t] (R from the quaternion
    qw qx qy qz), K from the camera model's fx, fy (or f), cx, cy, and the
    line "depth_min interval depth_num depth_max": the depths of the image's
    observed 3D points, their 1st percentile x 0.75 and 99th percentile
    x 1.25; depth_num = --max_d (192), or when 0 the inverse-depth count of
    one pixel's parallax at depth_min; interval = (max - min) /
    (depth_num - 1) / --interval_scale;
  - pair.txt: per image its --num_view (20) best partners by score, where
    the score of a pair is the number of 3D points both observe, or 0 when
    the 75th percentile of their triangulation angles is below 1 degree;
    the ranking (numpy's argsort, descending) runs over the whole score
    row, the image's own score 0 included, so an image with a zero-score
    partner can list itself; pair readers drop sources of score <= 0;
  - images/%08d.jpg: a .jpg copied, any other image re-encoded as JPEG
    (quality 95, with Pillow).
"""
from __future__ import annotations

import argparse
import os
import shutil
import struct
from typing import Dict, List, NamedTuple, Tuple

import numpy as np

# COLMAP camera models: id -> (name, parameter names in file order)
MODELS: Dict[int, Tuple[str, List[str]]] = {
    0: ("SIMPLE_PINHOLE", ["f", "cx", "cy"]),
    1: ("PINHOLE", ["fx", "fy", "cx", "cy"]),
    2: ("SIMPLE_RADIAL", ["f", "cx", "cy", "k"]),
    3: ("RADIAL", ["f", "cx", "cy", "k1", "k2"]),
    4: ("OPENCV", ["fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2"]),
    5: ("OPENCV_FISHEYE", ["fx", "fy", "cx", "cy", "k1", "k2", "k3", "k4"]),
    6: ("FULL_OPENCV", ["fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2", "k3", "k4", "k5", "k6"]),
    7: ("FOV", ["fx", "fy", "cx", "cy", "omega"]),
    8: ("SIMPLE_RADIAL_FISHEYE", ["f", "cx", "cy", "k"]),
    9: ("RADIAL_FISHEYE", ["f", "cx", "cy", "k1", "k2"]),
    10: ("THIN_PRISM_FISHEYE", ["fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2", "k3", "k4", "sx1", "sy1"]),
}
MODEL_IDS = {name: mid for mid, (name, _) in MODELS.items()}


class Camera(NamedTuple):
    id: int
    model: str
    width: int
    height: int
    params: np.ndarray


class Image(NamedTuple):
    id: int
    qvec: np.ndarray
    tvec: np.ndarray
    camera_id: int
    name: str
    point3d_ids: np.ndarray


# ---------------------------------------------------------------- reading

def _lines(path):
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                yield line


def read_cameras_text(path) -> Dict[int, Camera]:
    cams = {}
    for line in _lines(path):
        e = line.split()
        cams[int(e[0])] = Camera(int(e[0]), e[1], int(e[2]), int(e[3]), np.array(e[4:], np.float64))
    return cams


def read_images_text(path) -> Dict[int, Image]:
    """two lines per image; the second (its 2D points) is empty for an
    image without any"""
    with open(path) as f:
        lines = [ln.strip() for ln in f if not ln.startswith("#")]
    images = {}
    k = 0
    while k < len(lines):
        if not lines[k]:
            k += 1
            continue
        e = lines[k].split()
        pts = lines[k + 1].split() if k + 1 < len(lines) else []
        k += 2
        ids = np.array(pts[2::3], np.int64) if pts else np.zeros(0, np.int64)
        images[int(e[0])] = Image(int(e[0]), np.array(e[1:5], np.float64), np.array(e[5:8], np.float64), int(e[8]),
                                  e[9], ids)
    return images


def read_points3d_text(path) -> Dict[int, np.ndarray]:
    return {int(e[0]): np.array(e[1:4], np.float64) for e in (line.split() for line in _lines(path))}


class _Reader:
    def __init__(self, path):
        self.data = open(path, "rb").read()
        self.pos = 0

    def take(self, fmt):
        vals = struct.unpack_from("<" + fmt, self.data, self.pos)
        self.pos += struct.calcsize("<" + fmt)
        return vals

    def string(self):
        end = self.data.index(b"\0", self.pos)
        s = self.data[self.pos:end].decode()
        self.pos = end + 1
        return s


def read_cameras_binary(path) -> Dict[int, Camera]:
    r = _Reader(path)
    cams = {}
    for _ in range(r.take("Q")[0]):
        cid, mid, w, h = r.take("iiQQ")
        name, names = MODELS[mid]
        cams[cid] = Camera(cid, name, int(w), int(h), np.array(r.take("d" * len(names)), np.float64))
    return cams


def read_images_binary(path) -> Dict[int, Image]:
    r = _Reader(path)
    images = {}
    for _ in range(r.take("Q")[0]):
        iid = r.take("i")[0]
        q = np.array(r.take("dddd"), np.float64)
        t = np.array(r.take("ddd"), np.float64)
        cid = r.take("i")[0]
        name = r.string()
        n = r.take("Q")[0]
        ids = np.array(r.take("ddq" * n)[2::3], np.int64) if n else np.zeros(0, np.int64)
        images[iid] = Image(iid, q, t, cid, name, ids)
    return images


def read_points3d_binary(path) -> Dict[int, np.ndarray]:
    r = _Reader(path)
    pts = {}
    for _ in range(r.take("Q")[0]):
        pid = r.take("Q")[0]
        xyz = np.array(r.take("ddd"), np.float64)
        r.take("BBBd")
        n = r.take("Q")[0]
        r.take("ii" * n)
        pts[int(pid)] = xyz
    return pts


def read_model(folder, ext):
    if ext == ".bin":
        return (read_cameras_binary(os.path.join(folder, "cameras.bin")),
                read_images_binary(os.path.join(folder, "images.bin")),
                read_points3d_binary(os.path.join(folder, "points3D.bin")))
    return (read_cameras_text(os.path.join(folder, "cameras.txt")),
            read_images_text(os.path.join(folder, "images.txt")),
            read_points3d_text(os.path.join(folder, "points3D.txt")))


# ------------------------------------------------------------- conversion

def quat_to_rot(q) -> np.ndarray:
    """Rotation of the unit quaternion (w, x, y, z)."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def intrinsic(cam: Camera) -> np.ndarray:
    p = dict(zip(MODELS[MODEL_IDS[cam.model]][1], cam.params))
    fx = p.get("fx", p.get("f"))
    fy = p.get("fy", p.get("f"))
    return np.array([[fx, 0.0, p["cx"]], [0.0, fy, p["cy"]], [0.0, 0.0, 1.0]])


def depth_range(E, K, points3d, ids, max_d, interval_scale):
    """(depth_min, interval, depth_num, depth_max) of one image"""
    ids = [int(i) for i in ids if i != -1]
    zs = sorted(float((E @ np.append(points3d[i], 1.0))[2]) for i in ids)
    dmin = zs[int(len(zs) * 0.01)] * 0.75
    dmax = zs[int(len(zs) * 0.99)] * 1.25
    if max_d == 0:
        # the depth step at which one pixel of parallax appears at depth_min
        Ki, Rt = np.linalg.inv(K), E[:3, :3].T
        P1 = Rt @ (Ki @ np.array([K[0, 2], K[1, 2], 1.0]) * dmin - E[:3, 3])
        P2 = Rt @ (Ki @ np.array([K[0, 2] + 1.0, K[1, 2], 1.0]) * dmin - E[:3, 3])
        num = (1 / dmin - 1 / dmax) / (1 / dmin - 1 / (dmin + np.linalg.norm(P2 - P1)))
    else:
        num = max_d
    return dmin, (dmax - dmin) / (num - 1) / interval_scale, num, dmax


def pair_score(ids_i, ids_j, center_i, center_j, points3d) -> float:
    """shared observed 3D points, 0 when the 75th-percentile triangulation
    angle is below 1 degree"""
    in_j = set(int(x) for x in ids_j)
    angles = []
    for pid in ids_i:
        pid = int(pid)
        if pid == -1 or pid not in in_j:
            continue
        a, b = center_i - points3d[pid], center_j - points3d[pid]
        c = np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)
        angles.append(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))
    if not angles:
        return 0.0
    return 0.0 if sorted(angles)[int(len(angles) * 0.75)] < 1.0 else float(len(angles))


def convert(dense_folder, save_folder, model_ext=".txt", max_d=192, interval_scale=1.0, num_view=20):
    cams, images, points3d = read_model(os.path.join(dense_folder, "sparse"), model_ext)
    order = sorted(images)
    n = len(order)
    E, K, centers = [], [], []
    for iid in order:
        im = images[iid]
        e = np.eye(4)
        e[:3, :3] = quat_to_rot(im.qvec)
        e[:3, 3] = im.tvec
        E.append(e)
        K.append(intrinsic(cams[im.camera_id]))
        centers.append(-e[:3, :3].T @ e[:3, 3])
    score = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            score[i, j] = score[j, i] = pair_score(images[order[i]].point3d_ids, images[order[j]].point3d_ids,
                                                   centers[i], centers[j], points3d)
    cam_dir = os.path.join(save_folder, "cams")
    img_dir = os.path.join(save_folder, "images")
    for d in (cam_dir, img_dir):
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
    for i, iid in enumerate(order):
        dmin, interval, num, dmax = depth_range(E[i], K[i], points3d, images[iid].point3d_ids, max_d, interval_scale)
        with open(os.path.join(cam_dir, "%08d_cam.txt" % i), "w") as f:
            f.write("extrinsic\n")
            for r in range(4):
                f.write(" ".join(repr(float(v)) for v in E[i][r]) + "\n")
            f.write("\nintrinsic\n")
            for r in range(3):
                f.write(" ".join(repr(float(v)) for v in K[i][r]) + "\n")
            f.write("\n%f %f %f %f\n" % (dmin, interval, num, dmax))
    nv = min(num_view, n - 1)
    with open(os.path.join(save_folder, "pair.txt"), "w") as f:
        f.write("%d\n" % n)
        for i in range(n):
            best = np.argsort(score[i])[::-1][:nv]
            f.write("%d\n%d " % (i, nv) + " ".join("%d %d" % (k, score[i, k]) for k in best) + "\n")
    for i, iid in enumerate(order):
        src = os.path.join(dense_folder, "images", images[iid].name)
        dst = os.path.join(img_dir, "%08d.jpg" % i)
        if src.endswith(".jpg"):
            shutil.copyfile(src, dst)
        else:
            from PIL import Image as PILImage

            PILImage.open(src).convert("RGB").save(dst, "JPEG", quality=95)
    return n


def main():
    ap = argparse.ArgumentParser(description="COLMAP sparse model -> images/, cams/, pair.txt")
    ap.add_argument("--dense_folder", required=True)
    ap.add_argument("--save_folder", required=True)
    ap.add_argument("--max_d", type=int, default=192)
    ap.add_argument("--interval_scale", type=float, default=1.0)
    ap.add_argument("--num_view", type=int, default=20)
    ap.add_argument("--model_ext", default=".txt", choices=[".txt", ".bin"])
    a = ap.parse_args()
    os.makedirs(a.save_folder, exist_ok=True)
    n = convert(a.dense_folder, a.save_folder, a.model_ext, a.max_d, a.interval_scale, a.num_view)
    print(f"converted {n} images into {a.save_folder}")


if __name__ == "__main__":
    main()
