"""Host prior construction against independent checks:
  - support points: brute force over the 5x5 blocks;
  - Delaunay: every triangle's circumcircle is empty of other points (exact
    integer test), the triangles tile the convex hull (area), the triangles
    equal scipy.spatial.Delaunay's (Qhull) for points in general position,
    and on degenerate sets (lattices, co-circular points, collinear rows) a
    brute force over all triples with the index tie-break;
  - rasterisation: brute-force replay of the reference's float stepping;
  - plane fit: the numpy SVD null vector of [X 1];
  - KNN supplement: neighbours against a brute-force sort by
    (distance, index) and the plane through the first non-collinear triple.
"""
import ctypes

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay


def test_support_points_brute_force(hpm):
    rng = np.random.default_rng(0)
    H, W = 37, 53
    cost = rng.uniform(0, 0.3, (H, W)).astype(np.float32)
    cost[rng.uniform(size=(H, W)) < 0.2] = 2.0
    pts = hpm.support_points(cost)
    exp = []
    for col in range(0, W, 5):
        for row in range(0, H, 5):
            best, bp = np.float32(2.0), None
            for c in range(col, min(W, col + 5)):
                for r in range(row, min(H, row + 5)):
                    v = cost[r, c]
                    if v < 2.0 and best > v:
                        best, bp = v, (c, r)
            if best < np.float32(0.1):
                exp.append(bp)
    assert [tuple(p) for p in pts] == exp


def _incircle_exact(a, b, c, d):
    m = [[int(a[0] - d[0]), int(a[1] - d[1])], [int(b[0] - d[0]), int(b[1] - d[1])],
         [int(c[0] - d[0]), int(c[1] - d[1])]]
    sq = [x * x + y * y for x, y in m]
    return (m[0][0] * (m[1][1] * sq[2] - sq[1] * m[2][1]) - m[0][1] * (m[1][0] * sq[2] - sq[1] * m[2][0])
            + sq[0] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


@pytest.mark.parametrize("kind", ["random", "lattice"])
def test_delaunay_empty_circumcircle(hpm, kind):
    rng = np.random.default_rng(1)
    if kind == "random":
        pts = np.unique(rng.integers(0, 400, (600, 2)), axis=0).astype(np.int32)
        rng.shuffle(pts)
    else:  # support-point-like lattice with gaps: many cocircular quads
        g = np.array([(x, y) for x in range(0, 200, 5) for y in range(0, 150, 5)], np.int32)
        pts = g[rng.uniform(size=len(g)) < 0.8]
    tris = hpm.delaunay(pts)
    assert len(tris) > 0
    P = pts.astype(np.int64)
    for t in tris:
        a, b, c = P[t[0]], P[t[1]], P[t[2]]
        orient = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        assert orient > 0  # positively oriented, non-degenerate
    # empty circumcircle, exact, against every point (vectorised in int64 is
    # safe here: coordinates < 512)
    A = P[tris[:, 0]][:, None, :] - P[None, :, :]
    B = P[tris[:, 1]][:, None, :] - P[None, :, :]
    Cc = P[tris[:, 2]][:, None, :] - P[None, :, :]
    sa, sb, sc = (A ** 2).sum(-1), (B ** 2).sum(-1), (Cc ** 2).sum(-1)
    det = (A[..., 0] * (B[..., 1] * sc - sb * Cc[..., 1]) - A[..., 1] * (B[..., 0] * sc - sb * Cc[..., 0])
           + sa * (B[..., 0] * Cc[..., 1] - B[..., 1] * Cc[..., 0]))
    assert np.all(det <= 0)
    # the triangles tile the convex hull
    area = np.abs(np.cross(P[tris[:, 1]] - P[tris[:, 0]], P[tris[:, 2]] - P[tris[:, 0]])).sum() / 2
    hull = ConvexHull(pts.astype(np.float64))
    assert area == pytest.approx(hull.volume, rel=1e-9)
    if kind == "random":
        # general position: the Delaunay triangulation is unique
        assert len(tris) == len(Delaunay(pts.astype(np.float64)).simplices)


def _orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _incircle_perturbed(P, ia, ib, ic, id_):
    """sign of the incircle determinant of rows (x, y, x^2 + y^2 + eps_i, 1)
    with eps_0 >> eps_1 >> ...: the exact determinant, else the first
    non-zero cofactor of the lifted column by ascending point index"""
    a, b, c, d = P[ia], P[ib], P[ic], P[id_]
    det = _incircle_exact(a, b, c, d)
    if det:
        return 1 if det > 0 else -1
    cof = {ia: _orient(b, c, d), ib: -_orient(a, c, d), ic: _orient(a, b, d), id_: -_orient(a, b, c)}
    for i in sorted(cof):
        if cof[i]:
            return 1 if cof[i] > 0 else -1
    return 0


def _brute_force_triangulation(P):
    """every positively oriented triple whose perturbed circumcircle holds
    no other point, as (smallest index first, positively oriented)"""
    n = len(P)
    out = set()
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                o = _orient(P[i], P[j], P[k])
                if o == 0:
                    continue
                a, b, c = (i, j, k) if o > 0 else (i, k, j)
                if all(_incircle_perturbed(P, a, b, c, d) < 0 for d in range(n) if d not in (a, b, c)):
                    out.add((a, b, c))
    return out


def _point_sets():
    rng = np.random.default_rng(7)
    yield "random", rng.integers(0, 60, (35, 2))
    yield "lattice", np.array([(x, y) for x in range(0, 30, 5) for y in range(0, 25, 5)])
    g = np.array([(x, y) for x in range(0, 40, 5) for y in range(0, 30, 5)])
    yield "lattice_gaps", g[rng.uniform(size=len(g)) < 0.7]
    # eight points on one circle, its centre, and two outside
    yield "cocircular", np.array([(10, 0), (-10, 0), (0, 10), (0, -10), (6, 8), (-6, 8), (6, -8), (-8, -6), (0, 0),
                                  (20, 3), (-3, 21)]) + 30
    yield "collinear_rows", np.array([(x, y) for y in (0, 7, 14) for x in range(0, 40, 4)] + [(13, 3), (21, 10)])


@pytest.mark.parametrize("case", [c for c, _ in _point_sets()])
def test_delaunay_brute_force_with_tie_break(hpm, case):
    """the triangulation equals the brute-force set of triangles with an
    empty (perturbed) circumcircle, on degenerate sets: lattices (co-circular
    quads everywhere), points on one circle, collinear rows; and it is in the
    canonical order: smallest vertex index first, groups by ascending first
    vertex"""
    P = np.array(dict(_point_sets())[case], np.int64)
    pts = P.astype(np.int32)
    tris = hpm.delaunay(pts)
    got = [tuple(int(v) for v in t) for t in tris]
    assert len(set(got)) == len(got)
    want = _brute_force_triangulation([tuple(p) for p in P])
    assert set(got) == want
    assert all(t[0] < t[1] and t[0] < t[2] for t in got)
    firsts = [t[0] for t in got]
    assert firsts == sorted(firsts)
    # a permutation of the input gives the same triangles up to the relabel
    # when the tie-break order is carried along (the perturbation follows the
    # index, so it is relabelled too): here a reversal
    rev = pts[::-1].copy()
    n = len(pts)
    got_rev = {tuple(sorted((n - 1 - v for v in t))) for t in hpm.delaunay(rev)}
    want_rev = {tuple(sorted(t)) for t in _brute_force_triangulation([tuple(p) for p in rev.astype(np.int64)])}
    assert {tuple(sorted(n - 1 - v for v in t)) for t in want_rev} == got_rev


def test_delaunay_general_position_equals_qhull(hpm):
    rng = np.random.default_rng(3)
    pts = np.unique(rng.integers(0, 5000, (3000, 2)), axis=0).astype(np.int32)
    rng.shuffle(pts)
    got = {tuple(sorted(int(v) for v in t)) for t in hpm.delaunay(pts)}
    want = {tuple(sorted(int(v) for v in t)) for t in Delaunay(pts.astype(np.float64)).simplices}
    assert got == want


def test_delaunay_duplicates_use_the_first(hpm):
    pts = np.array([[0, 0], [10, 0], [0, 10], [10, 0], [10, 10], [0, 0]], np.int32)
    tris = hpm.delaunay(pts)
    used = set(int(v) for t in tris for v in t)
    assert used == {0, 1, 2, 4}


def test_rasterize_brute_force(hpm):
    pts = np.array([[2, 3], [40, 10], [15, 30], [45, 33], [3, 29]], np.int32)
    tris = np.array([[0, 1, 2], [1, 3, 2], [0, 2, 4]], np.int32)
    W, H = 50, 36
    got = hpm.rasterize_triangles(pts, tris, W, H)
    exp = np.zeros((H, W), np.float32)
    for idx, t in enumerate(tris):
        (x1, y1), (x2, y2), (x3, y3) = pts[t[0]], pts[t[1]], pts[t[2]]
        L = [np.float32(np.sqrt(float((x1 - x2) ** 2 + (y1 - y2) ** 2))),
             np.float32(np.sqrt(float((x1 - x3) ** 2 + (y1 - y3) ** 2))),
             np.float32(np.sqrt(float((x2 - x3) ** 2 + (y2 - y3) ** 2)))]
        step = np.float32(1.0 / float(max(L)))
        p = np.float32(0)
        while p < 1.0:
            q = np.float32(0)
            while float(q) < 1.0 - float(p):
                x = int(float(np.float32(p * np.float32(x1) + q * np.float32(x2))) + (1.0 - float(p) - float(q)) * x3)
                y = int(float(np.float32(p * np.float32(y1) + q * np.float32(y2))) + (1.0 - float(p) - float(q)) * y3)
                exp[y, x] = idx + 1
                q = np.float32(q + step)
            p = np.float32(p + step)
    assert np.array_equal(got, exp)
    assert (got > 0).mean() > 0.3


def _cam(hpm):
    K = [300.0, 0, 160.0, 0, 310.0, 120.0, 0, 0, 1]
    return hpm.camera_from_arrays(K, np.eye(3), [0, 0, 0], 320, 240, 1.0, 10.0)


def test_plane_fit_matches_svd(hpm):
    cam = _cam(hpm)
    rng = np.random.default_rng(2)
    depth = rng.uniform(2, 8, (60, 80)).astype(np.float32)
    for factor in (1.0, 0.5, 0.25):
        for _ in range(20):
            p = rng.integers(0, [80, 60], (3, 2)).astype(np.int32)
            got = hpm.plane_from_points(cam, p, depth, factor)
            K = np.array(cam.K, np.float64).reshape(3, 3)
            X = []
            for x, y in p:
                d = float(depth[y, x])
                X.append([d * (x - K[0, 2] * factor) / (K[0, 0] * factor),
                          d * (y - K[1, 2] * factor) / (K[1, 1] * factor), d, 1.0])
            _, _, vt = np.linalg.svd(np.array(X))
            n = vt[-1]
            n = n / np.linalg.norm(n[:3]) * (1 if n[3] >= 0 else -1)
            if abs(np.linalg.det(np.array(X)[:, :3])) < 1e-6 and np.allclose(np.cross(
                    np.subtract(p[1], p[0]), np.subtract(p[2], p[0])), 0):
                continue
            assert np.allclose(got, n, atol=2e-4), (got, n)
            assert got[3] >= 0


def test_knn_supplement_brute_force(hpm):
    cam = _cam(hpm)
    rng = np.random.default_rng(3)
    W, H = 60, 45
    # a 3D plane: inverse depth is affine in the pixel coordinates
    depth = (1.0 / (0.25 + 0.0005 * np.arange(W)[None, :] + 0.0008 * np.arange(H)[:, None])).astype(np.float32)
    pts = np.array([(x, y) for x in range(0, W, 5) for y in range(0, H, 5)], np.int32)
    pts = pts[rng.uniform(size=len(pts)) < 0.3] + rng.integers(0, 5, (1, 2)).astype(np.int32)
    pts = np.clip(pts, 0, [W - 1, H - 1]).astype(np.int32)
    mask = np.zeros((H, W), np.float32)
    mask[10:20, 10:30] = 7.0  # covered by a triangle: skipped
    k = 6
    m, planes = hpm.knn_supplement(cam, pts, depth, 1.0, k, 1.0, 10.0, mask)
    order = [(a, b, c) for a in range(5) for b in range(a + 1, 5) for c in range(b + 1, 5)]
    order += [(a, b, 5) for a in range(5) for b in range(a + 1, 5)]
    checked = 0
    for i in range(0, H, 3):
        for j in range(0, W, 3):
            if mask[i, j] > 0:
                assert m[i, j] == mask[i, j]
                continue
            d2 = (pts[:, 0] - j) ** 2 + (pts[:, 1] - i) ** 2
            nb = sorted(range(len(pts)), key=lambda s: (d2[s], s))[:k]
            tri = None
            for a, b, c in order:
                A, B, Cc = pts[nb[a]], pts[nb[b]], pts[nb[c]]
                if (B[0] - A[0]) * (Cc[1] - A[1]) - (B[1] - A[1]) * (Cc[0] - A[0]) != 0:
                    tri = np.array([A, B, Cc])
                    break
            assert tri is not None
            exp = hpm.plane_from_points(cam, tri, depth, 1.0)
            assert m[i, j] == -233.0
            assert np.array_equal(planes[i, j], exp)
            # the supplement plane reproduces the (planar) depth map
            d = hpm.lib().hpm_depth_from_plane_factor(ctypes.byref(cam), hpm.fptr(exp), j, i, 1.0)
            assert d == pytest.approx(depth[i, j], rel=1e-3)
            checked += 1
    assert checked > 100


def test_multiscale_count(hpm):
    L = hpm.lib()
    assert L.hpm_multiscale_count(640, 480) == 0
    assert L.hpm_multiscale_count(1600, 1066) == 1
    assert L.hpm_multiscale_count(3840, 2160) == 2
    assert L.hpm_multiscale_count(1000, 1000) == 0
    assert L.hpm_multiscale_count(2001, 10) == 1  # 2001 / 2 = 1000 (integer halving)
    assert L.hpm_multiscale_count(2002, 10) == 2  # 1001, then 500
