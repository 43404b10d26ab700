"""Pins the oracle to what fixes each function independently of its own code:
the reference's known-answer tests (cited per test), the textbook SGM
recurrence computed by brute force, stereo_match's byte-state recurrence
(paths.rs), sorting, and hand-worked closed forms. The reference ships no
golden vector files; its tests generate every input, and these restate them.
"""
import math

import numpy as np
import pytest

from nx_stereo_amd import scenes
from oracle import oracle

# --------------------------------------------------------------------------------------
# Census (census.rs:95-151; nx_stereo.ex:491-509)


def test_census_bit_order_follows_the_neighbour_order():
    image = np.full((7, 9), 100, np.uint8)
    image[1, 7] = 98  # two rows up, three columns right of (3, 4): neighbour (-2, +3)
    assert int(oracle.census(image)[3, 4]) == 1 << (8 + 6)


def test_census_thresholds():
    image = np.full((7, 9), 100, np.uint8)
    image[3, 4] = 102
    assert int(oracle.census(image)[3, 4]) == (1 << 48) - 1, "every neighbour two levels below"
    image[3, 4] = 101
    assert int(oracle.census(image)[3, 4]) == 0, "one level below does not count"


def test_census_clamps_at_the_border():
    # A bright corner pixel: every clamped neighbour reads the corner's own row/column.
    image = np.full((7, 9), 50, np.uint8)
    image[0, 0] = 60
    f = int(oracle.census(image)[0, 0])
    # Neighbours with dy < 0 clamp to row 0; with dx < 0 to column 0. A neighbour reads the
    # centre itself (not below it) iff both clamp: dy in -3..-1 and dx in -4..-1 -> 12 bits clear.
    assert bin(f).count("1") == 48 - 12


# --------------------------------------------------------------------------------------
# Costs (reference.rs:41-63)


def test_costs_closed_forms():
    rng = np.random.default_rng(0)
    left = rng.integers(0, 256, (9, 40), dtype=np.uint8)
    fl = oracle.census(left)
    c = oracle.costs(fl, fl, 16)
    assert (c[:, :, 0] == 0).all(), "identical features cost 0 at disparity 0"
    # Left of the image the right feature is 0: the cost is the left feature's popcount.
    popcount = np.vectorize(lambda v: bin(int(v)).count("1"))(fl)
    for d in range(16):
        assert (c[:, :d, d] == popcount[:, :d]).all()
    # A shift of the same texture by s costs 0 at d = s wherever both features see the same pixels.
    s = 5
    wide = rng.integers(0, 256, (9, 60), dtype=np.uint8)
    l, r = wide[:, :50], wide[:, s:50 + s]
    cs = oracle.costs(oracle.census(l), oracle.census(r), 16)
    assert (cs[:, s + 4:50 - 8, s] == 0).all()


# --------------------------------------------------------------------------------------
# Path recurrence: the textbook penalty-matrix form (Hirschmueller 2008, eq. 13), brute force.


def textbook_path(costs, dy, dx, p1, p2):
    h, w, d = costs.shape
    out = np.zeros((h, w, d), np.int64)
    disp = np.arange(d)
    penalty = np.where(disp[:, None] == disp[None, :], 0,
                       np.where(np.abs(disp[:, None] - disp[None, :]) == 1, p1, p2))
    ys = range(h - 1, -1, -1) if dy < 0 else range(h)
    xs = list(range(w - 1, -1, -1) if dx < 0 else range(w))
    for y in ys:
        for x in xs:
            py, px = y - dy, x - dx
            if 0 <= py < h and 0 <= px < w:
                prev = out[py, px]
                best = (prev[None, :] + penalty).min(axis=1) - prev.min()
            else:
                best = np.zeros(d, np.int64)
            out[y, x] = costs[y, x] + best
    return out


@pytest.mark.parametrize("p1,p2", [(0, 1), (8, 64), (78, 79), (3, 5)])
@pytest.mark.parametrize("direction", [(0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1)])
def test_path_equals_the_textbook_recurrence(p1, p2, direction):
    rng = np.random.default_rng(p1 * 100 + p2)
    costs = rng.integers(0, 49, (6, 11, 8), dtype=np.uint8)
    got = oracle.path(costs, *direction, p1, p2)
    want = textbook_path(costs.astype(np.int64), *direction, p1, p2)
    assert (got == want).all()
    assert got.max() <= p2 + 48


def byte_step(previous, previous_minimum, costs, p1, p2):
    """paths.rs:19-42 on bytes with 255 sentinels (its own formulation)."""
    d = len(costs)
    padded = np.concatenate([[255], previous, [255]]).astype(np.int64)
    jump = previous_minimum + p2
    out = np.empty(d, np.int64)
    for k in range(d):
        neighbour = min(min(padded[k], padded[k + 2]) + p1, 255)
        out[k] = min(padded[k + 1], jump, neighbour) - previous_minimum + costs[k]
    return out, int(out.min())


@pytest.mark.parametrize("p1,p2", [(0, 1), (8, 64), (78, 79)])
def test_path_equals_stereo_match_byte_states(p1, p2):
    """paths.rs:249-277: 200 random steps along one row, at the bounds of the penalties."""
    random = scenes.Random(9)
    costs = np.array([[[random.next() % 49 for _ in range(64)] for _ in range(200)]], np.uint8)
    got = oracle.path(costs, 0, 1, p1, p2)[0]
    state, minimum = np.zeros(64, np.int64), 0
    for t in range(200):
        state, minimum = byte_step(state, minimum, costs[0, t], p1, p2)
        assert (got[t] == state).all()
        assert state.max() <= p2 + 48


def test_a_constant_cost_path_keeps_the_cost():
    costs = np.full((5, 7, 16), 9, np.uint8)
    for direction in [(0, 1), (1, 0), (1, -1)]:
        assert (oracle.path(costs, *direction, 8, 64) == 9).all()


def test_aggregate_sums_the_directions():
    rng = np.random.default_rng(3)
    costs = rng.integers(0, 49, (5, 9, 8), dtype=np.uint8)
    dirs8 = [(0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1)]
    for paths, dirs in ((3, dirs8[:3]), (8, dirs8)):
        want = sum(textbook_path(costs.astype(np.int64), *d, 8, 64) for d in dirs)
        assert (oracle.aggregate(costs, paths, 8, 64) == want).all()


# --------------------------------------------------------------------------------------
# Selection (select.rs:207-261; reference.rs:148-186)


def select_at(sums, x, uniqueness=1.0):
    volume = np.zeros((1, x + 1, len(sums)), np.uint16)
    volume[0, x] = sums
    d, k = oracle.select(volume, uniqueness)
    return float(d[0, x]), int(k[0, x])


def test_select_ties_go_to_the_lower_disparity_and_the_edge_is_respected():
    sums = np.full(64, 100, np.uint16)
    sums[10] = sums[11] = 5
    assert select_at(sums, 63)[1] & 0xFFFF == 10
    sums[3] = 1
    assert select_at(sums, 63)[1] & 0xFFFF == 3
    sums[0] = 50
    assert select_at(sums, 2)[1] & 0xFFFF == 0, "disparities past x lose"


def test_select_subpixel_uniqueness_and_key_by_hand():
    sums = np.full(64, 100, np.uint16)
    sums[9], sums[10], sums[11] = 10, 4, 6
    # den = 10 + 6 - 8 = 8; correction = (10 - 6) / 16 = 0.25
    disparity, key = select_at(sums, 40, 1.0526316)
    assert disparity == 10.25 and key == (4 << 16) | 10
    assert select_at(sums, 10)[0] == 10.0, "no refinement at w = x"
    # Uniqueness: the second best outside w +- 1 must exceed min * u; an equal one fails.
    sums[30] = 4
    assert select_at(sums, 40, 1.0) == (-1.0, 0xFFFFFFFF)


def test_select_uniqueness_boundary():
    sums = np.full(64, 100, np.uint16)
    sums[10], sums[30] = 4, 5
    assert select_at(sums, 40, 1.2)[1] == (4 << 16) | 10  # 5 > 4.8
    assert select_at(sums, 40, 1.25)[1] == 0xFFFFFFFF     # 5 > 5.0 fails


def test_select_clamps_the_correction():
    sums = np.full(64, 100, np.uint16)
    # A tie at 9 and 10 goes to 9; lower = V[8] = 100, upper = V[10] = 4: den = 96,
    # correction = (100 - 4) / 192 = 0.5 exactly.
    sums[9], sums[10] = 4, 4
    assert select_at(sums, 40, 1.0) == (9.5, (4 << 16) | 9)
    sums[8] = 200  # (200 - 4) / (2 * 196) = 0.5 again; a larger gap clamps to +0.5
    sums[10] = 5
    # den = 200 + 5 - 8 = 197; (200 - 5) / 394 = 0.4949...
    assert abs(select_at(sums, 40, 1.0)[0] - (9 + np.float32(195) / np.float32(394))) < 1e-6


def test_select_matches_a_sorted_brute_force():
    rng = np.random.default_rng(11)
    for case in range(300):
        spread = (3, 20, 1016)[case % 3]
        sums = rng.integers(0, spread, 64).astype(np.uint16)
        x = int(rng.integers(0, 80))
        for u in (1.0, 1.0526316, 1.2):
            got = select_at(sums, x, u)
            values = sums.astype(np.int64) + np.where(np.arange(64) > x, 0x8000, 0)
            order = np.argsort(values, kind="stable")
            w, m = int(order[0]), int(values[order[0]])
            second = min(int(v) for d, v in enumerate(values) if abs(d - w) > 1)
            accepted = np.float32(second) > np.float32(m) * np.float32(u)
            if not accepted:
                assert got == (-1.0, 0xFFFFFFFF)
            else:
                assert got[1] == (m << 16) | w
                assert abs(got[0] - w) <= 0.5


# --------------------------------------------------------------------------------------
# Left-right check and median (check.rs:93-145; reference.rs:188-254)


def test_a_tie_on_a_right_pixel_favours_the_larger_disparity():
    keys = np.full((1, 16), 0xFFFFFFFF, np.uint32)
    keys[0, 10] = (3 << 16) | 5
    keys[0, 12] = (3 << 16) | 7
    disparity = np.full((1, 16), -1.0, np.float32)
    disparity[0, 10], disparity[0, 12] = 5.0, 7.0
    out = oracle.consistency(disparity, keys, 64)
    assert (out[0, 10], out[0, 12]) == (-1.0, 7.0)


def test_consistency_scatter_equals_stacked_minimum():
    """check.rs:113-131: the row scatter-min and the D-wide stacked minimum agree."""
    random = scenes.Random(8)
    width = 90
    for _ in range(200):
        keys = np.full((1, width), 0xFFFFFFFF, np.uint32)
        disparity = np.full((1, width), -1.0, np.float32)
        for x in range(width):
            if random.next() % 5 != 0:
                winner = random.next() % min(x + 1, 64)
                cost = random.next() % 4
                keys[0, x] = (cost << 16) | winner
                disparity[0, x] = max(winner + [0.0, 0.3, -0.4, 0.5][random.next() % 4], 0.0)
        right = np.full(width, 0xFFFFFFFF, np.uint64)
        for x in range(width):
            k = int(keys[0, x])
            if k != 0xFFFFFFFF:
                w = k & 0xFFFF
                right[x - w] = min(int(right[x - w]), (k & 0xFFFF0000) | (65535 - w))
        want = np.full(width, -1.0, np.float32)
        for x in range(width):
            v = float(disparity[0, x])
            if v < 0:
                continue
            lo, hi = int(np.floor(v)), int(np.ceil(v))

            def disagrees(rx, d):
                return 0 <= rx < width and right[rx] != 0xFFFFFFFF and abs(65535 - (int(right[rx]) & 0xFFFF) - d) > 1

            if not (disagrees(x - lo, lo) and disagrees(x - hi, hi)):
                want[x] = disparity[0, x]
        assert (oracle.consistency(disparity, keys, 64)[0] == want).all()


def test_median_is_the_sorted_middle_with_replicated_borders():
    image = ((np.arange(21 * 9) * 37 + 11) % 53 / 4.0 - 1.0).astype(np.float32).reshape(9, 21)
    got = oracle.median3(image)
    padded = np.pad(image, 1, mode="edge")
    for y in range(9):
        for x in range(21):
            assert got[y, x] == np.sort(padded[y:y + 3, x:x + 3].reshape(-1))[4]


# --------------------------------------------------------------------------------------
# The whole matcher (lib.rs:630-644; nx_stereo_test.exs:58-75; full_range_test.exs:40-55)


def test_a_shifted_scene_measures_its_disparity():
    left, right = scenes.scene(96, 64, 11)
    d = oracle.sgm(left, right, 64, 8)
    x, y = 60, 20
    u = (x - 4.0 - 6.0 * y / 64) / (1.0 + 30.0 / 96)
    assert abs(d[y, x] - (x - u)) < 1.0


def shifted_pair(height, width, shift, seed):
    wide = np.random.default_rng(seed).integers(0, 256, (height, width + shift), dtype=np.uint8)
    return wide[:, :width].copy(), wide[:, shift:shift + width].copy()


def test_recovers_the_disparity_of_a_shifted_texture():
    # The reference draws its texture from Erlang's :rand; this one is numpy's, seed 2.
    # Every interior pixel rounds to the shift, and the reference's 0.1 bound holds here.
    left, right = shifted_pair(24, 96, 5, 2)
    d = oracle.sgm(left, right)
    interior = d[4:20, 70:92]
    assert (np.round(interior) == 5.0).all()
    assert (np.abs(interior - 5.0) < 0.1).all()


def test_marks_pixels_without_a_reliable_match():
    left, _ = shifted_pair(16, 80, 0, 1)
    _, right = shifted_pair(16, 80, 0, 2)
    assert (oracle.sgm(left, right, uniqueness=1000.0) == -1.0).all()


@pytest.mark.parametrize("paths", [3, 8])
def test_recovers_a_shift_above_64_disparities(paths):
    wide = np.random.default_rng(14).integers(0, 256, (24, 256), dtype=np.uint8)
    left, right = wide[:, :160].copy(), wide[:, 96:256].copy()
    d = oracle.sgm(left, right, 128, paths)
    assert (np.round(d[4:20, 104:152]) == 96).all()


def test_a_match_at_disparity_0_reads_as_unmatched():
    """segment.rs:641-678: the stage maps d > 0 else -1."""
    left = scenes.noise(32, 24, 4)
    right = left.copy()
    right[12:] = np.concatenate([left[12:, 3:], np.repeat(left[12:, -1:], 3, 1)], 1)
    d = oracle.sgm(left, right, 64, 3)
    assert not (d == 0.0).any()
    assert (d[2:10, 8:24] == -1.0).all()


# --------------------------------------------------------------------------------------
# The CPU baseline is stereo_match's algorithm: it must equal the oracle (lib.rs:535-628).


@pytest.mark.parametrize("size", [(64, 48), (97, 31)])
@pytest.mark.parametrize("disparities", [64, 128, 192, 256])
@pytest.mark.parametrize("paths", [3, 8])
def test_cpu_baseline_matches_the_oracle(size, disparities, paths):
    w, h = size
    left, right = scenes.scene(w, h, 7)
    want = oracle.sgm(left, right, disparities, paths).view(np.uint32)
    for threads in (1, 3, 4):
        got = oracle.cpu_sgm(left, right, disparities, paths, threads=threads).view(np.uint32)
        assert (got == want).all()


def test_cpu_baseline_matches_the_oracle_on_a_planar_pair_at_256():
    """The checker of the large-pair GPU test (1920x1080x256 is beyond the oracle's full
    i32 volumes): a planar pair with disparities up to 248, where every range of the 256
    is in use."""
    left, right, _ = scenes.planar_pair(300, 40, seed=9, low=4, high=248)
    want = oracle.sgm(left, right, 256, 8).view(np.uint32)
    for threads in (0, 1, 5):
        assert (oracle.cpu_sgm(left, right, 256, 8, threads=threads).view(np.uint32) == want).all()


@pytest.mark.parametrize("threads", [1, 2, 7, 16])
def test_cpu_baseline_ring_and_chunks(threads):
    """3 paths through a ring of 3 rows (600 x 256 bytes a row: each slot refilled 13 times),
    and 8 paths over up to 37 column chunks whose neighbours hold each other back."""
    left, right, _ = scenes.planar_pair(600, 40, seed=4, low=4, high=200)
    for paths in (3, 8):
        want = oracle.sgm(left, right, 256, paths).view(np.uint32)
        got = oracle.cpu_sgm(left, right, 256, paths, threads=threads).view(np.uint32)
        assert (got == want).all(), (paths, threads)


def test_cpu_rectification_despeckle_and_depth_equal_the_oracle():
    """The baseline's map-based Rectification (the map built once per camera and kept) and
    its row-parallel SpeckleFilter, against the oracle's, over two calls (the second reusing
    the maps) and a changed camera (a new map)."""
    from nx_stereo_amd import stages

    cal = scenes.rig(200, 120)
    cams = {side: scenes.camera(**cal[side]) for side in ("left", "right")}
    frames = [scenes.noise(200, 120, 1), scenes.noise(200, 120, 2)]
    for offset in (0.0, 0.0, 1.5):
        right = dict(cams["right"], inverse_pr=stages.move_rows(cams["right"]["inverse_pr"], offset))
        for size in ((120, 200), (60, 100)):
            got = oracle.cpu_rectify_pair(frames[0], frames[1], cams["left"], right, size, threads=5)
            for image, frame, cam in zip(got, frames, (cams["left"], right)):
                want = oracle.rectify(frame, cam["k"], cam["d"], cam["inverse_pr"], size)
                assert np.array_equal(image, want), (offset, size)
    rng = np.random.default_rng(3)
    d = np.where(rng.random((53, 71)) < 0.3, -1.0, rng.uniform(0.5, 20.0, (53, 71))).astype(np.float32)
    for f, b in ((873.0, 0.2), (400.0, 0.06)):
        want_mm, want_m = oracle.depth(d, f, b)
        got_mm, got_m = oracle.cpu_depth(d, f, b, threads=4)
        assert np.array_equal(got_mm, want_mm) and np.array_equal(got_m.view(np.uint32),
                                                                    want_m.view(np.uint32))
    for radius, support, diff in ((3, 12, 1.0), (1, 3, 2.5), (5, 30, 0.5)):
        want = oracle.despeckle(d, support, diff, radius)
        for threads in (1, 6):
            got = oracle.cpu_despeckle(d, support, diff, radius, threads=threads)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (radius, threads)


def test_cpu_baseline_other_penalties_and_sizes():
    left, right = scenes.scene(80, 40, 3)
    for p1, p2, u in [(0, 1, 1.0), (4, 40, 1.2), (78, 79, 2.5), (10, 79, 1.0526316)]:
        for paths in (3, 8):
            want = oracle.sgm(left, right, 64, paths, p1, p2, u).view(np.uint32)
            assert (oracle.cpu_sgm(left, right, 64, paths, p1, p2, u).view(np.uint32) == want).all()
    for w, h in [(96, 64), (64, 48), (96, 64)]:
        left, right = scenes.scene(w, h, 5)
        assert (oracle.cpu_sgm(left, right, 64, 8).view(np.uint32)
                == oracle.sgm(left, right, 64, 8).view(np.uint32)).all()


# --------------------------------------------------------------------------------------
# Depth and the point cloud (depth_test.exs:6-98; depth.rs:95-154)


def test_depth_known_answers():
    mm, m = oracle.depth(np.array([[10.0, 0.0, -1.0, 0.0001, 0.36622]], np.float32), 400.0, 0.06)
    assert mm.tolist() == [[2400, 0, 0, 0, 65534]]
    assert float(m[0, 0]) == 2.4000000953674316
    assert (m[0, 1], m[0, 2]) == (0.0, 0.0)
    assert abs(m[0, 3] - 240000.0) <= 1.0
    mm, _ = oracle.depth(np.array([[10.0, 20.0]], np.float32), 400.0, -0.06)
    assert mm.tolist() == [[0, 0]], "a negative depth has no millimetres"
    mm, m = oracle.depth(np.array([10.0, 0.0, -1.0, 1.0e-4], np.float32), 100.0, 0.1)
    assert m[:3].tolist() == [1.0, 0.0, 0.0] and mm.tolist() == [1000, 0, 0, 0]
    mm, _ = oracle.depth(np.array([1.0, 0.9999], np.float32), 65535.0, 0.001)
    assert mm.tolist() == [65535, 0]


def test_cloud_known_answers():
    points, kept = oracle.cloud(np.full((48, 64), 2.0, np.float32), 4, 100.0, 32.0, 24.0)
    assert points.shape == (12, 16, 3) and kept == 12 * 16
    x, y, z = points[1, 1]
    assert abs(x - (4 - 32) * 2.0 / 100) < 1e-6 and abs(y - (4 - 24) * 2.0 / 100) < 1e-6 and z == 2.0
    depth = np.array([[0.0, 0.04, 0.06, 19.0, 21.0, 5.0]], np.float32)
    points, kept = oracle.cloud(depth, 1, 100.0, 0.0, 0.0)
    assert np.isnan(points[0, :, 2]).tolist() == [True, True, False, False, True, False] and kept == 3
    points, kept = oracle.cloud(np.array([[2.0, 9.0, 30.0, 9.0], [9.0, 9.0, 0.0, 9.0]], np.float32),
                                2, 2.0, 1.0, 0.0)
    assert kept == 1 and points[0, 0].tolist() == [-1.0, 0.0, 2.0]
    assert np.isnan(points[0, 1]).all()
    depth = (1.0 + np.arange(11, dtype=np.float32)[:, None] * np.ones((1, 8), np.float32))
    points, kept = oracle.cloud(depth, 4, 2.0, 0.0, 0.0)
    assert kept == 4 and points[1].reshape(-1).tolist() == [0.0, 10.0, 5.0, 10.0, 10.0, 5.0]
    depth = (np.arange(270, dtype=np.float32)[:, None] + np.ones((1, 8), np.float32))
    points, _ = oracle.cloud(depth, 4, 100.0, 0.0, 0.0, farthest=1000.0)
    assert points.shape == (67, 2, 3)
    assert [int(v - 1) for v in points[:, 0, 2]] == [v * 4 for v in range(67)]


# --------------------------------------------------------------------------------------
# Filters (filter_test.exs:7-100; filters.rs:330-458)


def speckled():
    d = np.full((12, 12), 10.0, np.float32)
    d[:, 8:] = -1.0
    d[5, 3] = 30.0
    d[6, 10] = 12.0
    return d


def test_despeckle_drops_islands_and_keeps_surfaces():
    out = oracle.despeckle(speckled(), 12, 1.0)
    assert out[5, 3] == -1.0 and out[6, 10] == -1.0
    assert out[0, 0] == 10.0 and out[11, 7] == 10.0 and out[0, 9] == -1.0
    assert (oracle.despeckle(speckled(), 0, 1.0) == speckled()).all()


def test_despeckle_counts_neighbours_within_max_diff_and_radius():
    image = np.full((9, 9), -1.0, np.float32)
    image[3:6, 3:6] = 29.5
    image[4, 4] = 30.0
    kept = lambda s, m, r=3: float(oracle.despeckle(image, s, m, r)[4, 4])
    assert kept(8, 1.0) == 30.0 and kept(9, 1.0) == -1.0 and kept(1, 0.25) == -1.0
    assert kept(8, 1.0, 1) == 30.0
    flat = np.full((5, 5), 10.0, np.float32)
    assert oracle.despeckle(flat, 9.0, 0.0, 1)[2, 2] == -1.0
    assert oracle.despeckle(flat, 8.0, 0.0, 1)[2, 2] == 10.0
    assert oracle.despeckle(flat, 24.0, 0.0, 3)[2, 2] == 10.0
    zero = np.zeros((10, 12), np.float32)
    zero[5, 5] = 0.5
    assert (oracle.despeckle(zero, 12.0, 1.0) == -1.0).all()


def test_temporal_known_answers():
    t = oracle.Temporal((3,))
    assert t.filter(np.array([10.0, 10.0, 10.0]), 0.5, 2.0, 1.0).tolist() == [10.0, 10.0, 10.0]
    assert t.filter(np.array([11.0, 20.0, -1.0]), 0.5, 2.0, 1.0).tolist() == [10.5, 20.0, 10.0]
    assert t.filter(np.array([11.0, 20.0, -1.0]), 0.5, 2.0, 1.0).tolist() == [10.75, 20.0, -1.0]
    t = oracle.Temporal((2,))
    t.filter(np.array([10.0, 0.0]), 0.4, 2.0, 1.0)
    assert t.filter(np.array([0.0, 1.0]), 0.4, 2.0, 1.0).tolist() == [10.0, 1.0]
    for cap, want in ((255.0, [10.0, 10.0, -1.0, -1.0]), (1.0, [10.0, 10.0, 10.0, 10.0])):
        t = oracle.Temporal((1,), cap)
        t.filter(np.array([10.0]), 1.0, 2.0, 2.0)
        assert [float(t.filter(np.array([-1.0]), 1.0, 2.0, 2.0)[0]) for _ in range(4)] == want
    t = oracle.Temporal((4,))
    t.previous[:] = [10.0, 10.0, -1.0, -1.0]
    t.missed[:] = [0, 0, 3, 0]
    assert t.filter(np.array([12.0, 20.0, 5.0, -1.0]), 0.5, 2.0, 0.0).tolist() == [11.0, 20.0, 5.0, -1.0]
    assert t.missed.tolist() == [0.0, 0.0, 0.0, 1.0]


def test_quality_known_answers():
    q = oracle.Quality((2, 2))
    assert q.sums(np.array([[2.0, 0.0], [-1.0, 4.0]])).tolist() == [2.0, 6.0, 20.0, 0.0, 0.0]
    assert q.sums(np.array([[3.0, 1.0], [1.0, 4.0]])).tolist() == [4.0, 9.0, 27.0, 2.0, 1.0]


@pytest.mark.parametrize("kernel_size", [3, 5])
def test_median_filter_is_the_sorted_middle_with_replicated_borders(kernel_size):
    """filters.rs:331-357's image, which holds negatives and zeros, at both window sizes."""
    image = ((np.arange(21 * 9) * 37 + 11) % 53 / 4.0 - 1.0).astype(np.float32).reshape(9, 21)
    got = oracle.median(image, kernel_size)
    r = kernel_size // 2
    padded = np.pad(image, r, mode="edge")
    for y in range(9):
        for x in range(21):
            window = np.sort(padded[y:y + kernel_size, x:x + kernel_size].reshape(-1))
            assert got[y, x] == window[kernel_size * kernel_size // 2]
    if kernel_size == 3:
        assert np.array_equal(got, oracle.median3(image))


# --------------------------------------------------------------------------------------
# Sharpness and RowAlignment sums (measures.rs:168-239; diagnostics_test.exs:17-95)


def test_sharpness_sums_neighbour_differences_and_grey():
    image = np.array([[10, 13, 13], [1, 1, 5]], np.uint8)
    assert oracle.sharpness_sums(image, image).tolist() == [25.0, 25.0, 25.0, 25.0, 4.0, 4.0]


def test_sharpness_counts_only_neighbours_both_above_black():
    image = np.array([[10, 13, 13], [0, 0, 4]], np.uint8)
    got = oracle.sharpness_sums(image, np.zeros((2, 3), np.uint8))
    assert got.tolist() == [9.0, 0.0, 23.0, 0.0, 2.0, 0.0]


def test_sharpness_diagnostics_cases():
    left = np.array([[10, 20, 10], [10, 10, 10]], np.uint8)
    right = np.full((2, 3), 10, np.uint8)
    # Left: differences 10, -10, 0, 0 over greys 10, 20, 10, 10; right: flat.
    assert oracle.sharpness_sums(left, right).tolist() == [200.0, 0.0, 50.0, 40.0, 4.0, 4.0]
    # Only the pair (10, 20) counts.
    one_row = np.array([[0, 200, 0, 10, 20]], np.uint8)
    got = oracle.sharpness_sums(one_row, np.zeros((1, 5), np.uint8))
    assert got.tolist() == [100.0, 0.0, 10.0, 0.0, 1.0, 0.0]
    # A single column has no pairs.
    column = np.array([[10], [20], [30]], np.uint8)
    assert oracle.sharpness_sums(column, column).tolist() == [0.0] * 6


def test_alignment_is_lowest_at_the_offset_the_right_image_is_moved_by():
    width, height, bands, reach, row_step = 32, 24, 4, 2, 4
    def texture(x, y):
        return (x * 31 + y * 17 + x * y * 7) % 251
    ys, xs = np.mgrid[0:height, 0:width]
    left = texture(xs, ys + 4).astype(np.uint8)
    right = texture(xs + 3, ys + 3).astype(np.uint8)  # rows one lower, three columns left
    sums = oracle.alignment_sums(left, right, np.full((height, width), 3.0, np.float32),
                                 bands, reach, row_step)
    assert sums.shape == (bands, 2 * reach + 2)
    for band in sums[1:]:
        assert int(np.argmin(band[:2 * reach + 1])) == reach + 1
        assert band[2 * reach + 1] == width // bands * height // row_step


def test_alignment_counts_no_pixel_at_disparity_0():
    image = np.full((24, 32), 90, np.uint8)
    sums = oracle.alignment_sums(image, image, np.zeros((24, 32), np.float32))
    assert sums.tolist() == [[0.0] * 6] * 4


def test_alignment_takes_its_bands_reach_and_rows_as_configured():
    image = np.full((6, 8), 90, np.uint8)
    sums = oracle.alignment_sums(image, image, np.ones((6, 8), np.float32), 2, 1, 2)
    assert sums.reshape(-1).tolist() == [0.0, 0.0, 0.0, 12.0, 0.0, 0.0, 0.0, 12.0]


@pytest.mark.parametrize("bands", [4, 8])
def test_alignment_has_no_samples_when_narrower_than_its_bands(bands):
    image = np.array([[10, 20, 30]], np.uint8)
    sums = oracle.alignment_sums(image, image, np.ones((1, 3), np.float32), bands)
    assert sums.shape == (bands, 6) and not sums.any()


def test_alignment_by_brute_force():
    """Every sum against a direct count on a random scene (rounding half away from zero,
    clamped columns and rows)."""
    rng = np.random.default_rng(5)
    h, w = 13, 29
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    right = rng.integers(0, 256, (h, w), dtype=np.uint8)
    disparity = np.where(rng.random((h, w)) < 0.3, -1.0, rng.uniform(0.0, 40.0, (h, w)))
    disparity[0, :4] = [0.5, 1.5, 2.5, 35.5]
    disparity = disparity.astype(np.float32)
    bands, reach, row_step = 3, 2, 3
    want = np.zeros((bands, 2 * reach + 2))
    bw = w // bands
    for y in range(0, h, row_step):
        for x in range(bands * bw):
            v = float(disparity[y, x])
            if not v > 0:
                continue
            sx = min(max(x - np.floor(v + 0.5), 0), w - 1)
            for o in range(2 * reach + 1):
                sy = min(max(y + o - reach, 0), h - 1)
                want[x // bw, o] += abs(int(right[sy, int(sx)]) - int(left[y, x]))
            want[x // bw, 2 * reach + 1] += 1
    got = oracle.alignment_sums(left, right, disparity, bands, reach, row_step)
    assert np.array_equal(got, want.astype(np.float32))


# --------------------------------------------------------------------------------------
# Rectification (rectify.rs:173-211; rectify_test.exs:39-162)

K = [100.0, 0.0, 32.0, 0.0, 100.0, 24.0, 0.0, 0.0, 1.0]
IDENTITY = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]


def projection(fx, cx, cy):
    return [fx, 0.0, cx, 0.0, 0.0, fx, cy, 0.0, 0.0, 0.0, 1.0, 0.0]


def rectify(image, cam, size=None):
    return oracle.rectify(image, cam["k"], cam["d"], cam["inverse_pr"], size or image.shape)


def test_identity_calibration_leaves_the_image_unchanged():
    image = np.random.default_rng(1).integers(0, 256, (48, 64), dtype=np.uint8)
    cam = scenes.camera(K, [], IDENTITY, projection(100.0, 32.0, 24.0))
    assert (rectify(image, cam) == image).all()
    frame = (np.arange(16 * 8) * 7 % 256).astype(np.uint8).reshape(8, 16)
    ident = {"k": np.array([100, 0, 8, 0, 100, 4, 0, 0, 1], np.float32), "d": np.zeros(5, np.float32),
             "inverse_pr": np.array([0.01, 0, -0.08, 0, 0.01, -0.04, 0, 0, 1], np.float32)}
    assert (rectify(frame, ident) == frame).all()


def test_a_principal_point_shift_moves_the_image():
    image = np.random.default_rng(2).integers(0, 256, (48, 64), dtype=np.uint8)
    cam = scenes.camera(K, [0.0] * 5, IDENTITY, projection(100.0, 35.0, 24.0))
    out = rectify(image, cam)
    assert (out[:, 3:64] == image[:, 0:61]).all()
    assert (out[:, 0:3] == 0).all()


def test_radial_distortion_maps_a_pixel_where_the_formula_says():
    k1, (u, v) = 0.5, (60, 40)
    x, y = (u - 32) / 100, (v - 24) / 100
    radial = 1 + k1 * (x * x + y * y)
    sx, sy = 100 * x * radial + 32, 100 * y * radial + 24
    image = np.zeros((48, 64), np.uint8)
    image[int(np.floor(sy)):int(np.floor(sy)) + 2, int(np.floor(sx)):int(np.floor(sx)) + 2] = 255
    cam = scenes.camera(K, [k1, 0, 0, 0, 0], IDENTITY, projection(100.0, 32.0, 24.0))
    assert rectify(image, cam)[v, u] == 255
    plain = scenes.camera(K, [], IDENTITY, projection(100.0, 32.0, 24.0))
    assert rectify(image, plain)[v, u] == 0


def test_halving_the_size_averages_each_square_of_four():
    width, height = 8, 4
    frame = (np.arange(2 * width * 2 * height) % 2 * 100).astype(np.uint8).reshape(2 * height, 2 * width)
    ident = {"k": np.array([100, 0, 4, 0, 100, 2, 0, 0, 1], np.float32), "d": np.zeros(5, np.float32),
             "inverse_pr": np.array([0.01, 0, -0.04, 0, 0.01, -0.02, 0, 0, 1], np.float32)}
    out = oracle.rectify(frame, ident["k"], ident["d"], ident["inverse_pr"], (height, width))
    assert out[1, 1] == 50


def test_maps_sample_where_the_projection_says():
    # maps at [10, 20] = (17, 10): a principal point at 35 moves the sample 3 px left. A frame
    # holding its own column index at row 10 reads back 17 there.
    cam = scenes.camera(K, [], IDENTITY, projection(100.0, 35.0, 24.0))
    frame = np.tile(np.arange(64, dtype=np.uint8), (48, 1))
    assert rectify(frame, cam)[10, 20] == 17
    rows = np.tile(np.arange(48, dtype=np.uint8)[:, None], (1, 64))
    assert rectify(rows, cam)[10, 20] == 10


def test_moving_rows_samples_that_many_rows_further_down():
    """stage/rectification.ex:84-89 (right_row_offset): column 2 of inverse(P3 R) gains
    rows x column 1, so output row v samples frame row v + rows. With an identity camera
    that is a plain shift by whole rows (0 past the frame), and half a row averages two."""
    from nx_stereo_amd import stages

    image = np.random.default_rng(5).integers(0, 256, (48, 64), dtype=np.uint8)
    cam = scenes.camera(K, [], IDENTITY, projection(100.0, 32.0, 24.0))
    for rows in (3, -2, 7):
        moved = dict(cam, inverse_pr=stages.move_rows(cam["inverse_pr"], rows))
        out = rectify(image, moved)
        lo, hi = max(0, -rows), min(48, 48 - rows)
        assert (out[lo:hi] == image[lo + rows:hi + rows]).all(), rows
        assert (out[:lo] == 0).all() and (out[hi:] == 0).all(), rows
    stripes = np.repeat((np.arange(48) % 2 * 100).astype(np.uint8)[:, None], 64, 1)
    moved = dict(cam, inverse_pr=stages.move_rows(cam["inverse_pr"], 0.5))
    assert (rectify(stripes, moved)[:47] == 50).all()
    # Through the stage's own params: the right camera moves, the left does not.
    geometry = {"left_camera": cam, "right_camera": cam}
    params = stages.Rectification(right_row_offset=3.0).params(geometry)
    assert (rectify(image, params["right"])[:45] == image[3:]).all()
    assert (rectify(image, params["left"]) == image).all()


def test_camera_inverse_is_the_inverse_of_the_rotated_projection():
    """rectify.ex:34-50: inverse_pr = inverse(P3 R) in f64, cast to f32. For the rig's
    rotations about the vertical axis, inverse(K R) = R^T inverse(K) in closed form."""
    cal = scenes.rig(728, 544)
    for side, sign in (("left", 1.0), ("right", -1.0)):
        c = cal[side]
        cam = scenes.camera(c["k"], c["d"], c["r"], c["p"])
        t = math.radians(sign * 0.25)
        r = np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])
        f, cx, cy = 873.0, 364.0, 272.0
        k_inv = np.array([[1 / f, 0, -cx / f], [0, 1 / f, -cy / f], [0, 0, 1]])
        want = (r.T @ k_inv).astype(np.float32).reshape(-1)
        got = cam["inverse_pr"]
        assert got.dtype == np.float32
        # f64 cofactors against the closed form: the same f32 value, or one ulp from it.
        ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        ulps[(got == 0) & (want == 0)] = 0  # +0 and -0
        assert (ulps <= 1).all(), (side, got, want)
        assert np.allclose(np.asarray(got, np.float64).reshape(3, 3)
                           @ (np.array([[f, 0, cx], [0, f, cy], [0, 0, 1]]) @ r), np.eye(3), atol=1e-6)
