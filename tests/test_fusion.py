"""Fusion (RunFusion, main.cpp:502-633) on the CPU: the binary PLY layout of
hpm_write_ply (StoreColorPlyFileBinaryPointCloud, HPM.cpp:403-456), and the
oracle's sequential fusion pinned against an independent pure-Python
restatement on tiny scenes (float32 scalar arithmetic in the reference's
expression order; exp and acos from the shared functions, so the two agree
bit for bit)."""
import ctypes as C
import math

import numpy as np

from test_oracle_pins import _lib

F = np.float32


def test_ply_layout(hpm, tmp_path):
    pts = np.zeros(3, hpm.POINT_DTYPE)
    pts["x"] = [1.5, np.nan, -2.0]
    pts["y"] = [2.5, 1.0, np.inf]
    pts["z"] = [3.5, 1.0, 4.0]
    pts["nx"], pts["ny"], pts["nz"] = 0.0, 0.0, -1.0
    pts["r"], pts["g"], pts["b"] = [10, 20, 30], [11, 21, 31], [12, 22, 32]
    path = str(tmp_path / "p.ply")
    hpm.write_ply(path, pts)
    data = open(path, "rb").read()
    header = (b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
              b"property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty uchar red\n"
              b"property uchar green\nproperty uchar blue\nend_header\n")
    assert data.startswith(header)
    body = np.frombuffer(data[len(header):], dtype=np.dtype([("xyz", "<f4", 3), ("n", "<f4", 3), ("rgb", "u1", 3)]))
    assert len(body) == 3
    assert np.array_equal(body["xyz"][0], [1.5, 2.5, 3.5])
    assert np.array_equal(body["xyz"][1], [0, 0, 0]) and np.array_equal(body["xyz"][2], [0, 0, 0])  # non-finite -> 0
    assert np.array_equal(body["rgb"][:, 0], [10, 20, 30]) and np.array_equal(body["rgb"][:, 2], [12, 22, 32])


def _shared(cpu, fn, v):
    L = _lib(cpu)
    x = np.array([v], np.float32)
    y = np.empty(1, np.float32)
    L.hpm_oracle_fmath(fn, x.ctypes.data_as(C.POINTER(C.c_float)), y.ctypes.data_as(C.POINTER(C.c_float)), 1)
    return F(y[0])


def _world(c, x, y, d):
    K, R, t = [F(v) for v in c.K], [F(v) for v in c.R], [F(v) for v in c.t]
    px = F(F(d * F(F(x) - K[2])) / K[0])
    py = F(F(d * F(F(y) - K[5])) / K[4])
    pz = F(d)
    tx = F(F(F(R[0] * px) + F(R[3] * py)) + F(R[6] * pz))
    ty = F(F(F(R[1] * px) + F(R[4] * py)) + F(R[7] * pz))
    tz = F(F(F(R[2] * px) + F(R[5] * py)) + F(R[8] * pz))
    cx = F(-F(F(F(R[0] * t[0]) + F(R[3] * t[1])) + F(R[6] * t[2])))
    cy = F(-F(F(F(R[1] * t[0]) + F(R[4] * t[1])) + F(R[7] * t[2])))
    cz = F(-F(F(F(R[2] * t[0]) + F(R[5] * t[1])) + F(R[8] * t[2])))
    return F(tx + cx), F(ty + cy), F(tz + cz)


def _project(c, X):
    K, R, t = [F(v) for v in c.K], [F(v) for v in c.R], [F(v) for v in c.t]
    tx = F(F(F(F(R[0] * X[0]) + F(R[1] * X[1])) + F(R[2] * X[2])) + t[0])
    ty = F(F(F(F(R[3] * X[0]) + F(R[4] * X[1])) + F(R[5] * X[2])) + t[1])
    tz = F(F(F(F(R[6] * X[0]) + F(R[7] * X[1])) + F(R[8] * X[2])) + t[2])
    d = F(F(F(K[6] * tx) + F(K[7] * ty)) + F(K[8] * tz))
    u = F(F(F(F(K[0] * tx) + F(K[1] * ty)) + F(K[2] * tz)) / d)
    v = F(F(F(F(K[3] * tx) + F(K[4] * ty)) + F(K[5] * tz)) / d)
    return u, v, d


def _python_fusion(cpu, cams, depths, normals, colors, problems, acmm=False):
    """acmm: ACMM+NESP's RunFusion (ACMM+NESP/windows/main.cpp:164-300): means
    of the reference's and the consistent sources' points, normals (times the
    float reciprocal, cv::Vec3f /= float) and colours, 2 sources needed"""
    masks = [np.zeros(d.shape, np.uint8) for d in depths]
    out = []
    for i, (_, srcs) in enumerate(problems):
        rc = cams[i]
        used = [(-1, -1)] * len(srcs)
        for r in range(rc.height):
            for c in range(rc.width):
                if masks[i][r, c] == 1:
                    continue
                ref_depth = F(depths[i][r, c])
                if ref_depth <= 0:
                    continue
                X = _world(rc, c, r, ref_depth)
                n_cons, dyn = 0, F(0)
                sx = list(X)
                sn = [F(v) for v in normals[i][r, c]]
                scol = [F(v) for v in colors[i][r, c]]
                for j, s in enumerate(srcs):
                    sc = cams[s]
                    u, v, _ = _project(sc, X)
                    sr, sc_ = int(F(v + F(0.5))), int(F(u + F(0.5)))
                    if not (0 <= sc_ < sc.width and 0 <= sr < sc.height):
                        continue
                    if masks[s][sr, sc_] == 1:
                        continue
                    sd = F(depths[s][sr, sc_])
                    if sd <= 0:
                        continue
                    Y = _world(sc, sc_, sr, sd)
                    tu, tv, pd = _project(rc, Y)
                    dx, dy = float(F(F(c) - tu)), float(F(F(r) - tv))
                    reproj = F(math.sqrt(dx * dx + dy * dy))
                    rdd = F(abs(F(pd - ref_depth)) / ref_depth)
                    a, b = normals[i][r, c], normals[s][sr, sc_]
                    dot = F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))
                    ang = _shared(cpu, 3, dot)
                    ang = F(0) if ang != ang else ang
                    if reproj < F(2) and rdd < F(0.01) and ang < F(0.174533):
                        sx = [F(sx[k] + Y[k]) for k in range(3)]
                        sn = [F(sn[k] + F(b[k])) for k in range(3)]
                        scol = [F(scol[k] + F(colors[s][sr, sc_][k])) for k in range(3)]
                        used[j] = (sc_, sr)
                        idx = F(F(reproj + F(F(200) * rdd)) + F(ang * F(10)))
                        dyn = F(dyn + _shared(cpu, 0, -idx))
                        n_cons += 1
                if acmm and n_cons >= 2:
                    k = F(F(n_cons) + F(1))
                    ik = F(F(1) / k)
                    out.append((tuple(F(v / k) for v in sx), tuple(F(v * ik) for v in sn),
                                tuple(int(F(v / k)) for v in scol)))
                    for j, s in enumerate(srcs):
                        if used[j][0] != -1:
                            masks[s][used[j][1], used[j][0]] = 1
                elif not acmm and n_cons >= 1 and float(dyn) > 0.3 * n_cons:
                    out.append((X, tuple(normals[i][r, c]), tuple(colors[i][r, c])))
                    for j, s in enumerate(srcs):
                        if used[j][0] != -1:
                            masks[s][used[j][1], used[j][0]] = 1
    return out


def test_oracle_fusion_matches_python(hpm, cpu):
    cfg = hpm.scene_config(seed=8, width=40, height=30, fx=32.0, fy=32.0, cx=20.0, cy=15.0, num_views=3,
                           num_planes=3, num_boxes=2)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    rng = np.random.default_rng(0)
    # ground truth with noise, holes and a few flipped normals: every branch of the test is taken
    depths = [(d * (1 + 0.004 * rng.standard_normal(d.shape))).astype(np.float32) for d in gd]
    for d in depths:
        d[rng.random(d.shape) < 0.05] = 0.0
    normals = [n.astype(np.float32).copy() for n in gn]
    for n in normals:
        flip = rng.random(n.shape[:2]) < 0.05
        n[flip] = -n[flip]
    colors = [np.repeat(im.astype(np.uint8)[..., None], 3, axis=2) for im in imgs]
    problems = [(i, [j for j in range(3) if j != i]) for i in range(3)]
    got = cpu.fusion(cams, depths, normals, colors, problems, hpm.POINT_DTYPE)
    want = _python_fusion(cpu, cams, depths, normals, colors, problems)
    assert len(got) == len(want) and len(got) > 100
    for g, (X, n, col) in zip(got, want):
        assert (g["x"], g["y"], g["z"]) == X
        assert (g["nx"], g["ny"], g["nz"]) == n
        assert (g["r"], g["g"], g["b"]) == col


def test_oracle_acmm_fusion_matches_python(hpm, cpu):
    """ACMM+NESP's RunFusion (means over the consistent sources, 2 needed)
    in the oracle against the pure-Python loops, bit for bit"""
    cfg = hpm.scene_config(seed=9, width=40, height=30, fx=32.0, fy=32.0, cx=20.0, cy=15.0, num_views=4,
                           num_planes=3, num_boxes=2)
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    rng = np.random.default_rng(1)
    depths = [(d * (1 + 0.003 * rng.standard_normal(d.shape))).astype(np.float32) for d in gd]
    for d in depths:
        d[rng.random(d.shape) < 0.05] = 0.0
    normals = [n.astype(np.float32).copy() for n in gn]
    for n in normals:
        flip = rng.random(n.shape[:2]) < 0.05
        n[flip] = -n[flip]
    colors = [np.repeat(im.astype(np.uint8)[..., None], 3, axis=2) for im in imgs]
    problems = [(i, [j for j in range(4) if j != i]) for i in range(4)]
    got = cpu.fusion(cams, depths, normals, colors, problems, hpm.POINT_DTYPE, acmm=True)
    want = _python_fusion(cpu, cams, depths, normals, colors, problems, acmm=True)
    plain = cpu.fusion(cams, depths, normals, colors, problems, hpm.POINT_DTYPE)
    assert len(got) == len(want) and len(got) > 100 and len(got) != len(plain)
    for g, (X, n, col) in zip(got, want):
        assert (g["x"], g["y"], g["z"]) == X
        assert (g["nx"], g["ny"], g["nz"]) == n
        assert (g["r"], g["g"], g["b"]) == col
