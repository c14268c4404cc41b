"""Pins the oracle's propagation logic on hand-built micro-scenes.

The bilateral NCC is replaced by a cost hook (hpm_oracle_problem.cost_hook),
so each test builds the exact cost fields that the NESP scans, JudgeExtend,
the joint view selection, adoption and PlaneHypothesisRefinement act on,
and the expected outcome follows by hand from the text of HPM-MVS/HPM.cu:

  - every scan's chosen neighbour (HPM.cu:1251-1471), including the
    maximum kept by right_far (HPM.cu:1330), first-minimum ties, and cells
    outside the scan paths that must be ignored;
  - invalid directions near the border keep cost 0 except cost_array[0][0]
    = 2 (HPM.cu:1244) and win FindMinCostIndex as the last minimum
    (HPM.cu:29-41), which blocks adoption;
  - JudgeExtend's float thresholds (HPM.cu:1036-1059): a cost of exactly
    1.2f is not bad, a cost of exactly the good threshold is not good, the
    per-round threshold 0.8*exp(-iter^2*(3-k)/90), and the shifted scan of
    ExtendedLeftUpPropagation (HPM.cu:780-843);
  - view-selection priors, probabilities, CDF and the 15 draws
    (HPM.cu:1548-1613, 86-100), and the all-zero case whose NaN costs take
    the non-prior fallback (HPM.cu:1716-1721);
  - adoption: the last minimum of final_costs, the depth-range gate, the
    strict improvement, and the selected-view mask (HPM.cu:1615-1722);
  - the prior path: restricted costs, the argmax adoption, the view mask
    left alone with prior mask 0, and the refinement's acceptance against
    restricted_cost = 0 (HPM.cu:1652-1704, 1205-1210);
  - the five refinement candidates and their sequential acceptance
    (HPM.cu:1143-1220);
  - the three initialisation modes (HPM.cu:510-558) and the top-k initial
    cost with ties (HPM.cu:399-436).
"""
import math

import numpy as np
import pytest

F32 = np.float32


def _cam(hpm, W, H, tx, dmin, dmax, R=None):
    c = hpm.HpmCamera()
    c.K[:] = [20.0, 0.0, W / 2, 0.0, 20.0, H / 2, 0.0, 0.0, 1.0]
    c.R[:] = R if R is not None else [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    c.t[:] = [tx, 0.0, 0.0]
    c.width, c.height = W, H
    c.depth_min, c.depth_max = dmin, dmax
    return c


class Micro:
    """A W x H reference view with nsrc sources whose costs come from a table.

    Every pixel starts with its own fronto-parallel camera-frame plane
    (0, 0, -1, 3 + idx / 1024) (exact in FP32; its depth is the same value
    at every pixel), so a plane handed to the cost hook names the pixel it
    was read from. rows[idx] is that hypothesis's cost per source view
    (default 1.0); any other plane (refinement candidates) costs
    other(px, py, view, plane), 1.9 by default."""

    def __init__(self, hpm, cpu, W=32, H=32, nsrc=4, dmin=1.0, dmax=10.0, R=None, depths=False):
        self.W, self.H, self.nsrc = W, H, nsrc
        self.default = 1.0
        cams = [_cam(hpm, W, H, -0.2 * i, dmin, dmax, R) for i in range(nsrc + 1)]
        rng = np.random.default_rng(0)
        imgs = [rng.integers(0, 256, (H, W)).astype(F32) for _ in range(nsrc + 1)]
        p = hpm.default_params()
        p.num_images = nsrc + 1
        p.depth_min, p.depth_max = dmin, dmax
        self.params = p
        self.cams = cams
        dm = [np.full((H, W), 4.0, F32) for _ in range(nsrc + 1)] if depths else None
        self.o = cpu.Oracle(cams, imgs, p, depths=dm)
        idx = np.arange(W * H, dtype=np.float64).reshape(H, W)
        self.o.hyp[:] = 0.0
        self.o.hyp[..., 2] = -1.0
        self.o.hyp[..., 3] = (3.0 + idx / 1024.0).astype(F32)
        self.o.cost[:] = 1.0
        self.o.sel[:] = (1 << nsrc) - 1
        self.rows = {}
        self.other = lambda px, py, v, plane: 1.9
        self.calls = []
        self.o.set_cost_hook(self._hook)

    def idx(self, x, y):
        return y * self.W + x

    def plane(self, x, y):
        return self.o.hyp[y, x].copy()

    def key(self, plane):
        if plane[0] == 0.0 and plane[1] == 0.0 and plane[2] == -1.0:
            k = (float(plane[3]) - 3.0) * 1024.0
            if k == int(k) and 0 <= k < self.W * self.H:
                return int(k)
        return None

    def _hook(self, px, py, v, plane):
        k = self.key(plane)
        self.calls.append((px, py, v, k))
        if k is not None:
            row = self.rows.get(k)
            return float(row[v - 1]) if row is not None else self.default
        return float(self.other(px, py, v, plane))

    def set_row(self, x, y, row):
        self.rows[self.idx(x, y)] = np.asarray(row, F32)

    def run(self, x, y, it=0):
        colour = (x + y) & 1  # black: x + y even (HPM.cu:1729-1734)
        tr = self.o.trace_pixel(x, y)
        self.o.half_iteration(it, colour)
        assert tr.hit == 1
        return tr


@pytest.fixture
def micro(hpm, cpu):
    return lambda **kw: Micro(hpm, cpu, **kw)


def _xy(m, pos):
    return (pos % m.W, pos // m.W)


# ---------------------------------------------------------------- NESP scans

def test_scans_choose_the_neighbours_the_text_names(micro):
    """Pixel (16, 16) of 32 x 32: every direction valid and every scan step
    inside the image. The candidate cells of each scan, from HPM.cu:
      up_far    (16, 11), (16, 9), (16, 7), (16, 5), (16, 3)
      down_far  (16, 21), (16, 23), (16, 25), (16, 27), (16, 29)
      left_far  (11, 16), (9, 16), (7, 16), (5, 16), (3, 16)
      right_far (21, 16), (23, 16), (25, 16), (27, 16), (29, 16)  [maximum]
      left_up   (11, 10), (9, 10), (9, 8), (7, 8), (7, 6), (5, 6), (5, 4), (3, 4)
      right_up  (22, 11), (22, 9), (24, 9), (24, 7), (26, 7), (26, 5), (28, 5), (28, 3)
      right_down(21, 22), (23, 22), (23, 24), (25, 24), (25, 26), (27, 26), (27, 28), (29, 28)
      left_down (10, 21), (10, 23), (8, 23), (8, 25), (6, 25), (6, 27), (4, 27), (4, 29)
    Cells just off each path hold lower costs and must be ignored."""
    m = micro()
    c = m.o.cost
    # decoys off the paths (all of the other colour, all lower than anything on a path)
    for (x, y) in [(16, 1), (16, 31), (1, 16), (31, 16), (8, 7), (23, 10), (22, 23), (9, 24), (15, 12), (12, 15)]:
        c[y, x] = 0.0
    c[7, 16] = 0.5                      # up_far: the third cell
    c[23, 16] = 0.5                     # down_far: tie between (16, 23) and (16, 27) ...
    c[27, 16] = 0.5                     # ... the first in scan order stays (strict <)
    c[16, 3] = 0.4                      # left_far: the last cell
    c[16, 23] = 0.1                     # right_far keeps the maximum: a minimum is ignored ...
    c[16, 27] = 1.5                     # ... and the largest cost is chosen
    c[6, 7] = 0.3                       # left_up
    c[11, 22] = 0.2                     # right_up: the start cell itself
    c[28, 29] = 0.25                    # right_down: the last cell
    c[25, 6] = 0.35                     # left_down
    c[16, 16] = 1.0
    tr = m.run(16, 16)
    got = [_xy(m, p) for p in tr.pos_scan]
    want = [(7, 6), (16, 7), (22, 11), (16, 23), (29, 28), (3, 16), (6, 25), (27, 16)]
    assert got == want
    assert list(tr.flag) == [1] * 8


def test_scan_near_the_border_stops_or_skips(micro):
    """Pixel (7, 24) of 32 x 32 (x + y odd: red). left_up starts at (2, 18):
    its first step (0, 18) is inside, the second (0, 16) too, the third
    (-2, 16) leaves the image and the scan stops (break). left_down starts
    at (1, 29): its first step (1, 31) is inside, the second (-1, 31)
    leaves the image and x only decreases afterwards, so the missing break
    changes nothing. up_far at y = 24 scans rows 19, 17, 15, 13, 11;
    down_far scans row 29 and row 31 (24 < 32 - 5 - 2 holds for i = 1
    only)."""
    m = micro()
    c = m.o.cost
    c[18, 0] = 0.6     # left_up second cell
    c[16, 0] = 0.7     # left_up third cell (higher)
    c[29, 1] = 0.9     # left_down start
    c[31, 1] = 0.2     # left_down first step (y + 2 = 31 inside)
    c[11, 7] = 0.3     # up_far last cell
    c[31, 7] = 0.0     # down_far's second cell
    tr = m.run(7, 24)
    got = dict(zip(range(8), [_xy(m, p) for p in tr.pos_scan]))
    assert tr.flag[0] == 1 and got[0] == (0, 18)
    assert tr.flag[6] == 1 and got[6] == (1, 31)
    assert tr.flag[1] == 1 and got[1] == (7, 11)
    assert tr.flag[3] == 1 and got[3] == (7, 31)


def test_invalid_directions_keep_zero_costs_and_win_the_minimum(micro):
    """Pixel (2, 16): left_up (x > 4), left_far (x > 4) and left_down
    (x > 5) are invalid. cost_array = {2.0f} leaves their rows at 0 except
    [0][0] = 2, final_costs of rows 5 and 6 are 0, FindMinCostIndex keeps
    the last minimum (6), whose flag is false: nothing is adopted even
    though every valid neighbour is better than the current hypothesis."""
    m = micro()
    for x, y in [(2, 11), (2, 9), (2, 7), (2, 5), (2, 3), (2, 21), (2, 23), (2, 25), (2, 27), (2, 29)]:
        m.set_row(x, y, [0.3] * 4)
    m.set_row(2, 16, [0.9] * 4)
    tr = m.run(2, 16)
    assert list(tr.flag) == [0, 1, 1, 1, 1, 0, 0, 1]
    ca = np.array(tr.ca)[:, :4]
    assert ca[0, 0] == 2.0 and np.all(ca[0, 1:] == 0.0)
    assert np.all(ca[5] == 0.0) and np.all(ca[6] == 0.0)
    assert tr.fc[5] == 0.0 and tr.fc[6] == 0.0
    assert tr.min_idx == 6
    assert tr.adopted == -1
    # the cost is the re-scored current hypothesis (candidates cost 1.9)
    assert m.o.cost[16, 2] == np.float32(tr.cost_now)
    assert abs(tr.cost_now - 0.9) < 1e-6
    assert np.array_equal(m.o.hyp[16, 2], m.plane(2, 16))


# ------------------------------------------------------------- JudgeExtend

def _judge_case(micro, row, it=0):
    m = micro()
    # left_up of (16, 16) is its start cell (11, 10) when every cost is equal
    m.set_row(11, 10, row)
    tr = m.run(16, 16, it)
    assert _xy(m, tr.pos_scan[0]) == (11, 10)
    return tr


def test_judge_extend_bad_threshold_is_float_1_2(micro):
    """float bad_threshold = 1.2 (HPM.cu:1042): a cost equal to 1.2f is not
    > 1.2f. Three views at exactly 1.2f and one good view: good_sum = 1,
    bad_sum = 0, no extension. One float step above 1.2f makes all three
    bad (bad_sum = 3 > 2) and the scan extends."""
    b = F32(1.2)
    tr = _judge_case(micro, [b, b, b, 0.5])
    assert list(tr.judge) == [0, -1, -1]
    up = np.nextafter(b, F32(2))
    tr = _judge_case(micro, [up, up, up, 0.5])
    assert tr.judge[0] == 1


def test_judge_extend_good_threshold_is_strict(micro):
    """Iteration 0: good = 0.8 * exp(0) = 0.8f. A cost equal to it is not
    good (strict <): no good view, so the scan extends; one float step
    below it is good and stops."""
    g = F32(0.8)
    tr = _judge_case(micro, [g, 1.0, 1.0, 1.0])
    assert tr.judge[0] == 1
    tr = _judge_case(micro, [np.nextafter(g, F32(0)), 1.0, 1.0, 1.0])
    assert list(tr.judge) == [0, -1, -1]


def test_judge_extend_threshold_per_round(micro):
    """Iteration 1: good_k = 0.8 * exp(-(3 - k) / 90): 0.77378 (k = 0),
    0.78242 (k = 1), 0.79116 (k = 2). A best cost of 0.78 is not good in
    round 0 (extend; the round-0 scan repeats the first window, so the cost
    row is unchanged) but good in round 1 (stop)."""
    tr = _judge_case(micro, [0.78, 1.0, 1.0, 1.0], it=1)
    assert list(tr.judge) == [1, 0, -1]
    assert tr.pos[0] == 10 * 32 + 11


def test_extension_scans_the_shifted_window(micro):
    """Pixel (30, 30) of 48 x 48, iteration 0. The first left_up scan
    chooses (25, 24) (all path costs equal) whose row has no good view, so
    the scan extends. Round 0 rescans the same window; round 1 starts at
    center - 5 - 6W - 8 - 8W = (17, 16) and visits (15, 16), (15, 14),
    (13, 14), (13, 12), (11, 12), (11, 10), (9, 10): its minimum (13, 14)
    is lower than the stored cost at (25, 24), so it becomes left_up and
    its row replaces cost_array[0]; that row has a good view and round 2
    is not taken. The lower cost at (12, 13), off the path, is ignored."""
    m = micro(W=48, H=48)
    c = m.o.cost
    m.set_row(25, 24, [1.0, 1.0, 1.0, 1.0])
    c[14, 13] = 0.4
    c[13, 12] = 0.1
    m.set_row(13, 14, [0.3, 1.0, 1.0, 1.0])
    tr = m.run(30, 30)
    assert _xy(m, tr.pos_scan[0]) == (25, 24)
    assert list(tr.judge) == [1, 1, 0]
    assert _xy(m, tr.pos[0]) == (13, 14)
    assert np.allclose(np.array(tr.ca[0][:4]), [0.3, 1.0, 1.0, 1.0])


def test_extension_keeps_the_position_when_the_far_window_is_worse(micro):
    """Same pixel, but every cell of the round-1 window costs more than the
    stored cost at the current left_up: costMin > costs[position[0]], so
    symbol = false, the position and cost_array[0] stay (HPM.cu:818-830);
    the next round judges the same row again."""
    m = micro(W=48, H=48)
    c = m.o.cost
    c[24, 25] = 0.5
    for (x, y) in [(17, 16), (15, 16), (15, 14), (13, 14), (13, 12), (11, 12), (11, 10), (9, 10)]:
        c[y, x] = 0.7
    m.set_row(25, 24, [1.0, 1.0, 1.0, 1.0])
    tr = m.run(30, 30)
    assert list(tr.judge) == [1, 1, 1]
    assert _xy(m, tr.pos[0]) == (25, 24)


# ---------------------------------------------------------- view selection

def _view_selection_scene(micro, rows_by_dir, masks):
    """Pixel (16, 16): the chosen neighbour of direction d gets cost row
    rows_by_dir[d]; masks = selected-view masks of (16, 11), (16, 21),
    (11, 16), (21, 16)."""
    m = micro()
    cells = [(11, 10), (16, 11), (22, 11), (16, 21), (21, 22), (11, 16), (10, 21), (21, 16)]
    for d, (x, y) in enumerate(cells):
        m.o.cost[y, x] = 0.5  # the scan starts are chosen (ties keep the first)
        m.set_row(x, y, rows_by_dir[d])
    # right_far keeps the maximum: its start must be the largest on its path
    m.o.cost[16, 21] = 1.5
    for (x, y), mk in zip([(16, 11), (16, 21), (11, 16), (21, 16)], masks):
        m.o.sel[y, x] = mk
    m.set_row(16, 16, [0.95] * 4)
    tr = m.run(16, 16)
    assert [_xy(m, p) for p in tr.pos] == cells
    return m, tr


def test_view_selection_probabilities(micro):
    """Iteration 0: threshold 0.8. Per view (columns of cost_array):
      view 0: 3 costs below 0.8 (0.1, 0.2, 0.3), none above 1.2
              -> mean of exp(-c^2 / 0.18) over the three;
      view 1: 2 below -> count > 2 fails -> exp(-0.8^2 / 0.32) = e^-2;
      view 2: 4 below but 3 above 1.2 -> 0;
      view 3: none below, 2 above 1.2 -> e^-2.
    Priors from the four far neighbours' masks 0b0001, 0b0011, 0b0111,
    0b1111: 0.9 per set bit, 0.1 per clear one -> 3.6, 2.8, 2.0, 1.2."""
    v0 = [0.1, 0.2, 0.3, 0.9, 0.9, 0.9, 0.9, 0.9]
    v1 = [0.1, 0.2, 0.9, 0.9, 0.9, 0.9, 0.9, 0.9]
    v2 = [0.1, 0.2, 0.3, 0.4, 1.3, 1.3, 1.3, 0.9]
    v3 = [1.3, 1.3, 0.9, 0.9, 0.9, 0.9, 0.9, 0.9]
    rows = [[v0[d], v1[d], v2[d], v3[d]] for d in range(8)]
    m, tr = _view_selection_scene(micro, rows, [0b0001, 0b0011, 0b0111, 0b1111])
    prior = np.array(tr.vprior[:4], np.float64)
    assert np.allclose(prior, [3.6, 2.8, 2.0, 1.2], rtol=1e-6, atol=0)
    f = lambda c: math.exp(-c * c / 0.18)
    p0 = (f(0.1) + f(0.2) + f(0.3)) / 3
    want = np.array([p0 * 3.6, math.exp(-2.0) * 2.8, 0.0, math.exp(-2.0) * 1.2])
    pdf = np.array(tr.pdf[:4], np.float64)
    assert np.allclose(pdf, want, rtol=2e-6, atol=0)
    cdf = np.array(tr.cdf[:4], np.float64)
    assert np.allclose(cdf, np.cumsum(want) / want.sum(), rtol=2e-6, atol=0)
    assert cdf[2] == cdf[1]
    # 15 draws: each goes to the first view whose CDF exceeds it
    vw = np.zeros(4)
    for u in tr.draws:
        for v in range(4):
            if cdf[v] > u:
                vw[v] += 1
                break
    assert np.array_equal(np.array(tr.vw[:4]), vw)
    assert vw[2] == 0 and vw.sum() == 15
    assert tr.tsel == sum(1 << v for v in range(4) if vw[v] > 0)
    assert tr.wn == vw.sum()
    # final_costs: weighted mean of each direction's row over the drawn views
    fc = (np.array(rows, np.float64) * vw).sum(1) / vw.sum()
    assert np.allclose(np.array(tr.fc), fc, rtol=1e-6, atol=1e-7)
    # far neighbours invalid -> no prior contribution: the text adds priors only for flag[2i+1]


def test_all_zero_probabilities_take_the_nan_fallback(micro):
    """Every view has three costs above 1.2 in its column: all sampling
    probabilities are 0, the CDF is 0 * inf = NaN, no draw selects a view,
    weight_norm = 0, every final cost and the re-scored cost are 0 / 0 =
    NaN. FindMinCostIndex keeps index 0 (no comparison with NaN holds), the
    ordinary adoption fails (NaN < NaN is false), and the non-prior NaN
    fallback adopts left_up's hypothesis with an empty view mask while the
    stored cost stays NaN (it assigns cost_now, not costs[center])."""
    rows = [[1.3] * 4] * 3 + [[0.5] * 4] * 5
    m, tr = _view_selection_scene(micro, rows, [0b1111] * 4)
    assert np.all(np.array(tr.vw[:4]) == 0) and tr.wn == 0.0 and tr.tsel == 0
    assert np.all(np.isnan(np.array(tr.cdf[:4])))
    assert np.all(np.isnan(np.array(tr.fc)))
    assert math.isnan(tr.cost_now)
    assert tr.min_idx == 0
    assert tr.adopted == 0 and tr.nan_fallback == 1
    assert np.array_equal(m.o.hyp[16, 16], m.plane(11, 10))
    assert math.isnan(m.o.cost[16, 16])
    assert m.o.sel[16, 16] == 0
    assert not any(tr.cand_accepted)


# ----------------------------------------------------------------- adoption

def _adoption_scene(micro, dir_costs, cur=0.5, **kw):
    """Direction d's chosen neighbour costs dir_costs[d] on every view, so
    final_costs[d] = dir_costs[d] whatever views are drawn."""
    return _view_selection_scene(micro, [[c] * 4 for c in dir_costs], [0b1111] * 4)


def test_adoption_takes_the_last_minimum(micro):
    """final_costs 0.3 at right_up (2) and right_down (4): FindMinCostIndex
    keeps the last (4); it beats the current 0.95, lies in the depth range,
    and is adopted with its final cost and the drawn view mask."""
    m, tr = _adoption_scene(micro, [0.9, 0.8, 0.3, 0.7, 0.3, 0.6, 0.95, 0.9])
    assert tr.fc[2] == tr.fc[4]
    assert tr.min_idx == 4 and tr.adopted == 4
    assert np.array_equal(m.o.hyp[16, 16], m.plane(21, 22))
    assert m.o.cost[16, 16] == np.float32(tr.fc[4])
    assert m.o.sel[16, 16] == tr.tsel
    assert not any(tr.cand_accepted)  # every refinement candidate costs 1.9


def test_adoption_needs_a_strictly_lower_cost(micro):
    """The best neighbour costs exactly what the current hypothesis costs
    (0.95): final_costs[i] < cost_now fails, nothing is adopted."""
    m, tr = _adoption_scene(micro, [0.95] * 8)
    assert tr.fc[tr.min_idx] == np.float32(tr.cost_now)
    assert tr.adopted == -1
    assert np.array_equal(m.o.hyp[16, 16], m.plane(16, 16))


def test_adoption_gates_on_the_depth_range(micro):
    """The best neighbour's plane lies at depth 3 + idx/1024 >= 3.1; with
    depth_max 3.05 it is out of range and not adopted, the cost is the
    re-scored one."""
    rows = [[c] * 4 for c in [0.9, 0.8, 0.2, 0.7, 0.6, 0.6, 0.95, 0.9]]
    mm = micro(dmin=1.0, dmax=3.05)
    cells = [(11, 10), (16, 11), (22, 11), (16, 21), (21, 22), (11, 16), (10, 21), (21, 16)]
    for d, (x, y) in enumerate(cells):
        mm.o.cost[y, x] = 0.5
        mm.set_row(x, y, rows[d])
    mm.o.cost[16, 21] = 1.5
    mm.set_row(16, 16, [0.95] * 4)
    tr = mm.run(16, 16)
    assert tr.min_idx == 2 and mm.plane(22, 11)[3] > 3.05
    assert tr.adopted == -1
    assert mm.o.cost[16, 16] == np.float32(tr.cost_now)


# -------------------------------------------------------------- refinement

def test_refinement_candidates_and_sequential_acceptance(micro):
    """PlaneHypothesisRefinement without prior: candidates (depth, normal) =
    (d_rand, n_cur), (d_cur, n_rand), (d_rand, n_rand), (d_cur, n_pert),
    (d_pert, n_cur) with d_rand in [dmin, dmax], d_pert within 2% of the
    current depth, n_pert within the 0.02*pi Euler perturbation, each plane's
    d = -n . X(depth). A candidate replaces the hypothesis when its depth is
    in range and its cost beats the running cost (strict). Costs here grow
    with the distance to a target depth, so acceptance can be replayed."""
    m = micro()
    target = 4.5

    def depth_of(plane, px, py):
        fx, cx, cy = 20.0, 16.0, 16.0
        return -plane[3] * fx / ((px - cx) * plane[0] + (py - cy) * plane[1] + fx * plane[2])

    m.other = lambda px, py, v, pl: min(1.9, 0.05 + abs(depth_of(pl, px, py) - target) * 0.2 + 0.1 * (1 + pl[2]))
    m.default = 1.1  # every neighbour is worse than the current hypothesis (and not bad: < 1.2)
    m.set_row(16, 16, [0.9] * 4)
    cur = m.plane(16, 16).astype(np.float64)
    tr = m.run(16, 16)
    assert tr.adopted == -1
    cand = np.array(tr.cand, np.float64)
    cd = np.array(tr.cand_depth, np.float64)
    d_cur = cur[3]
    assert np.array_equal(cand[0, :3], cur[:3]) and np.array_equal(cand[4, :3], cur[:3])
    assert np.array_equal(cand[1, :3], cand[2, :3])
    assert cd[0] == cd[2] and 1.0 <= cd[0] <= 10.0
    assert cd[1] == cd[3] == np.float32(d_cur)
    assert abs(cd[4] / d_cur - 1) <= 0.02 + 1e-6
    for i in range(5):
        n = cand[i, :3]
        assert abs(np.linalg.norm(n) - 1) < 1e-6
        assert abs(depth_of(cand[i], 16, 16) - cd[i]) < 1e-5 * cd[i]
    ang = math.acos(min(1.0, float(np.dot(cand[3, :3], cur[:3]))))
    assert ang <= 1.5 * 0.02 * math.pi + 1e-6
    # replay the acceptance: running cost starts at the re-scored 0.9
    run = F32(tr.cost_now)
    final = cur.astype(F32)
    for i in range(5):
        c = F32(tr.cand_cost[i])
        assert np.isclose(c, m.other(16, 16, 1, cand[i]), rtol=1e-5)
        ok = 1.0 <= depth_of(cand[i], 16, 16) <= 10.0 and c < run
        assert tr.cand_accepted[i] == int(ok), i
        if ok:
            run, final = c, cand[i].astype(F32)
    assert np.array_equal(m.o.hyp[16, 16], final)
    assert m.o.cost[16, 16] == run


# -------------------------------------------------------------- prior path

def _prior_scene(micro, mask_val, dir_costs, prior_depth, cur_cost=0.6):
    m = micro()
    W, H = m.W, m.H
    planes = np.zeros((H, W, 4), F32)
    planes[..., 2] = -1.0
    planes[..., 3] = prior_depth
    mask = np.full((H, W), mask_val, np.uint32)
    m.o.set_prior(planes, mask)
    m.o.pb.params.planar_prior = 1
    cells = [(11, 10), (16, 11), (22, 11), (16, 21), (21, 22), (11, 16), (10, 21), (21, 16)]
    for d, (x, y) in enumerate(cells):
        m.o.cost[y, x] = 0.5
        m.set_row(x, y, [dir_costs[d]] * 4)
    m.o.cost[16, 21] = 1.5
    m.set_row(16, 16, [cur_cost] * 4)
    m.o.sel[16, 16] = 0b0110
    return m, cells


def test_prior_adoption_by_restricted_cost(micro):
    """Prior mask 1, prior plane fronto-parallel at depth 3.3. The
    restricted cost of a candidate with final cost c and depth d is
    exp(-c^2 / 0.18) * (0.5 + exp(-(d - 3.3)^2 / (2 s_d^2)) * exp(-0^2 /
    (2 s_a^2))), s_d = (10 - 1) / 64 (same normal: angle 0). left_up (depth
    3.323, cost 0.4) wins the argmax over down_far (depth 3.672, the lowest
    final cost 0.2); it is adopted with the drawn view mask, and
    restricted_cost becomes its value."""
    m, cells = _prior_scene(micro, 1, [0.4, 0.5, 0.5, 0.2, 0.5, 0.5, 0.5, 0.5], prior_depth=3.3)
    depths = [float(m.plane(x, y)[3]) for (x, y) in cells]
    d_cur = float(m.plane(16, 16)[3])
    tr = m.run(16, 16)
    sd = 9.0 / 64
    fc = np.array(tr.fc, np.float64)
    want = [math.exp(-fc[i] ** 2 / 0.18) * (0.5 + math.exp(-(depths[i] - 3.3) ** 2 / (2 * sd * sd)))
            for i in range(8)]
    assert np.allclose(np.array(tr.rfc), want, rtol=1e-5)
    best = int(np.argmax(want))
    assert best == 0 and tr.max_idx == best and int(np.argmin(fc)) == 3
    rcn = math.exp(-tr.cost_now ** 2 / 0.18) * (0.5 + math.exp(-(d_cur - 3.3) ** 2 / (2 * sd * sd)))
    assert abs(tr.rcn - rcn) < 1e-5 * rcn
    assert want[best] > rcn
    assert tr.adopted == best
    assert tr.restricted_cost == np.float32(tr.rfc[best])
    assert m.o.sel[16, 16] == tr.tsel


def test_prior_refinement_accepts_the_first_candidate_when_nothing_was_adopted(micro):
    """Prior mask 1 and no neighbour beats the current restricted cost:
    restricted_cost stays 0 (HPM.cu:1652), so the first in-range refinement
    candidate is accepted whatever its cost (its restricted cost is > 0),
    here one that costs 1.9 against the current 0.2."""
    m, cells = _prior_scene(micro, 1, [1.0] * 8, prior_depth=3.49, cur_cost=0.2)
    tr = m.run(16, 16)
    assert tr.adopted == -1 and tr.restricted_cost == 0.0
    assert np.isclose(tr.cand_cost[0], 1.9, rtol=1e-6)
    first = next(i for i in range(5) if 1.0 <= tr.cand_depth[i] <= 10.0)
    assert tr.cand_accepted[first] == 1
    assert tr.cand_rcost[first] > 0.0


def test_prior_mask_zero_adopts_without_the_view_mask(micro):
    """planar_prior on, mask 0: adoption by the minimum final cost as
    without the prior, but selected_views is not written (HPM.cu:1696-1704),
    unlike the path without the prior (HPM.cu:1706-1714)."""
    m, cells = _prior_scene(micro, 0, [0.9, 0.8, 0.3, 0.7, 0.9, 0.6, 0.95, 0.9], prior_depth=3.3, cur_cost=0.95)
    m.o.sel[16, 16] = 0b10000  # a bit no drawn mask can hold (4 sources)
    tr = m.run(16, 16)
    assert tr.max_idx == -1 and tr.min_idx == 2 and tr.adopted == 2
    assert np.array_equal(m.o.hyp[16, 16], m.plane(22, 11))
    assert tr.tsel != 0 and m.o.sel[16, 16] == 0b10000


# ----------------------------------------------------------- initialisation

def test_initial_cost_top_k_with_ties(micro):
    """ComputeMultiViewInitialCostandSelectedViews: costs (0.5, 0.2, 2.0,
    0.2): 3 valid (< 2), top_k = min(3, 4) = 3, cost = (0.2 + 0.2 + 0.5) / 3,
    mask = views with cost <= 0.5 (ties included) = 0b1011. All at 2:
    cost 2, mask 0."""
    m = micro()
    m.other = lambda px, py, v, pl: [0.5, 0.2, 2.0, 0.2][v - 1] if px != 3 else 2.0
    m.o.init()
    want = (F32(0.2) + F32(0.2) + F32(0.5)) / F32(3)
    assert m.o.cost[10, 10] == want
    assert m.o.sel[10, 10] == 0b1011
    assert m.o.cost[10, 3] == 2.0 and m.o.sel[10, 3] == 0


def test_random_initialisation(micro):
    """GenerateRandomPlaneHypothesis: depth uniform in [dmin, dmax], a unit
    normal facing the camera (n . ray < 0), d = -n . X(depth)."""
    m = micro()
    m.other = lambda px, py, v, pl: 0.5
    m.o.init()
    h = m.o.hyp.astype(np.float64)
    n = h[..., :3]
    assert np.allclose(np.linalg.norm(n, axis=-1), 1, atol=1e-6)
    ys, xs = np.mgrid[0:32, 0:32]
    ray = np.stack([(xs - 16) / 20.0, (ys - 16) / 20.0, np.ones_like(xs, np.float64)], -1)
    assert np.all((n * ray).sum(-1) <= 0)
    depth = -h[..., 3] * 20 / ((xs - 16) * n[..., 0] + (ys - 16) * n[..., 1] + 20 * n[..., 2])
    assert depth.min() >= 1.0 - 1e-4 and depth.max() <= 10.0 + 1e-4
    assert depth.max() - depth.min() > 5  # spread over the range
    assert np.all(m.o.cost == F32(0.5))


def test_prior_initialisation_modes(micro):
    """Prior mode (HPM.cu:529-553): where mask > 0 and the cost >= 0.1 the
    hypothesis is the prior normal perturbed by the 0.06*pi Euler angles
    with w = the prior's w perturbed by up to 6% (stored as is, not as a
    distance); elsewhere the stored hypothesis keeps its normal (no
    TransformNormal2RefCam, although the reference camera is rotated) and
    w becomes -n . X(w as depth)."""
    c, s = math.cos(0.3), math.sin(0.3)
    m = micro(R=[c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c])
    nb = np.array([0.28, -0.15, -0.948])
    m.o.hyp[..., :3] = (nb / np.linalg.norm(nb)).astype(F32)
    m.other = lambda px, py, v, pl: 0.5
    planes = np.zeros((32, 32, 4), F32)
    planes[..., 0] = 0.6
    planes[..., 2] = -0.8
    planes[..., 3] = 4.0
    mask = np.zeros((32, 32), np.uint32)
    mask[:, 16:] = 1
    m.o.cost[:] = 0.5
    m.o.cost[20:, :] = 0.05  # below 0.1: keep
    before = m.o.hyp.copy()
    m.o.set_prior(planes, mask)
    m.o.pb.params.planar_prior = 1
    m.o.init()
    h = m.o.hyp
    a = h[:20, 16:]
    assert np.all(a[..., 3] >= 0.94 * 4.0 - 1e-6) and np.all(a[..., 3] <= 1.06 * 4.0 + 1e-6)
    cosang = a[..., 0] * 0.6 + a[..., 2] * -0.8
    assert np.all(np.arccos(np.clip(cosang, -1, 1)) <= 1.5 * 0.06 * math.pi + 1e-5)
    assert np.all(np.abs(np.linalg.norm(a[..., :3], axis=-1) - 1) < 1e-6)
    for (y0, y1, x0, x1) in [(0, 20, 0, 16), (20, 32, 0, 32)]:
        b = h[y0:y1, x0:x1]
        assert np.array_equal(b[..., :3], before[y0:y1, x0:x1, :3])
        ys, xs = np.mgrid[y0:y1, x0:x1]
        d = before[y0:y1, x0:x1, 3].astype(np.float64)
        X = np.stack([d * (xs - 16) / 20, d * (ys - 16) / 20, d], -1)
        want = -(before[y0:y1, x0:x1, :3].astype(np.float64) * X).sum(-1)
        assert np.allclose(b[..., 3], want, rtol=1e-6)


def test_geometric_initialisation_rotates_the_normal(micro, hpm):
    """Geometric-consistency mode (HPM.cu:523-528): the stored (world
    normal, depth) becomes (R n, -(R n) . X(depth))."""
    c, s = math.cos(0.3), math.sin(0.3)
    R = [c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c]
    m = micro(R=R, depths=True)
    m.other = lambda px, py, v, pl: 0.5
    h0 = np.zeros((32, 32, 4), F32)
    h0[..., 0] = 0.0
    h0[..., 2] = -1.0
    h0[..., 3] = 4.0
    m.o.hyp[:] = h0
    m.o.pb.params.geom_consistency = 1
    m.o.init()
    Rm = np.array(R, np.float64).reshape(3, 3)
    n = Rm @ np.array([0.0, 0.0, -1.0])
    h = m.o.hyp.astype(np.float64)
    assert np.allclose(h[..., :3], n, atol=1e-7)
    ys, xs = np.mgrid[0:32, 0:32]
    X = np.stack([4.0 * (xs - 16) / 20, 4.0 * (ys - 16) / 20, 4.0 * np.ones_like(xs, np.float64)], -1)
    assert np.allclose(h[..., 3], -(X * n).sum(-1), rtol=1e-6, atol=1e-6)
