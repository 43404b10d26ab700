"""The oracle's band search and coarse-to-fine (nx_stereo.ex:164-471), each function checked
against a route other than itself: the full-range functions pinned in test_oracle.py, scipy
routines, closed forms and the reference's own band tests (nx_stereo_test.exs:115-225)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle
from tests.band_scenes import interior, layered, pair

DIRECTIONS = [(0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1)]


# ---- shrink ---------------------------------------------------------------------------------

def test_shrink_closed_forms():
    assert (oracle.shrink(np.full((6, 8), 77, np.uint8)) == 77).all()
    block = np.array([[0, 0], [0, 2]], np.uint8)  # (2 + 2) // 4 = 1: the +2 rounds
    assert oracle.shrink(block)[0, 0] == 1
    assert oracle.shrink(np.array([[0, 0], [0, 1]], np.uint8))[0, 0] == 0
    assert oracle.shrink(np.full((2, 2), 255, np.uint8))[0, 0] == 255
    image = np.zeros((4, 6), np.uint8)
    image[2:4, 4:6] = 200  # only the last block of the second block row
    out = oracle.shrink(image)
    assert out.shape == (2, 3)
    assert out[1, 2] == 200 and out.sum() == 200


def test_shrink_is_the_rounded_block_mean():
    image = np.random.default_rng(3).integers(0, 256, (14, 22), dtype=np.uint8)
    sums = image.astype(np.int64).reshape(7, 2, 11, 2).sum(axis=(1, 3))
    assert (oracle.shrink(image) == (sums + 2) // 4).all()


# ---- census windows -------------------------------------------------------------------------

def test_sparse_window_is_the_census():
    image = np.random.default_rng(4).integers(0, 256, (12, 20), dtype=np.uint8)
    assert (oracle.census_window(image, "sparse_9x7") == oracle.census(image)).all()


def brute_dense(image, anchor):
    h, w = image.shape
    padded = np.pad(image.astype(np.int64), ((2, 2), (4, 4)), mode="edge")
    out = np.zeros((h, w), np.uint64)
    bit = 0
    for dy in range(-2, 3):
        for dx in range(-4, 5):
            if dy == 0 and dx == 0:
                continue
            neighbour = padded[2 + dy:2 + dy + h, 4 + dx:4 + dx + w]
            out |= (neighbour + 1 < anchor).astype(np.uint64) << np.uint64(bit)
            bit += 1
    assert bit == 44
    return out


def test_dense_windows_against_a_brute_force():
    image = np.random.default_rng(5).integers(0, 256, (11, 23), dtype=np.uint8)
    assert (oracle.census_window(image, "dense_9x5") == brute_dense(image, image)).all()
    gauss = np.outer([1, 2, 1], [1, 2, 1])
    weighted = ndimage.correlate(image.astype(np.int64), gauss, mode="nearest")
    anchor = (weighted + 16 * image.astype(np.int64) + 24) // 32
    assert (oracle.census_window(image, "dense_9x5_blend") == brute_dense(image, anchor)).all()


def test_dense_windows_closed_forms():
    flat = np.full((9, 12), 90, np.uint8)
    for window in ("dense_9x5", "dense_9x5_blend"):
        assert (oracle.census_window(flat, window) == 0).all()
    # A bright pixel over a flat background: every neighbour in its window is below it.
    spot = flat.copy()
    spot[4, 6] = 200
    dense = oracle.census_window(spot, "dense_9x5")[4, 6]
    assert int(dense) == (1 << 44) - 1
    # Blended: (4*200 + 12*90 + 16*200 + 24) // 32 = 159 > 91, so still every bit.
    assert int(oracle.census_window(spot, "dense_9x5_blend")[4, 6]) == (1 << 44) - 1
    # At 92 the blended anchor is (20*92 + 12*90 + 24) // 32 = 92: 90 + 1 < 92 holds;
    # at 91 it is (20*91 + 1080 + 24) // 32 = 91 and 90 + 1 < 91 does not.
    for centre, bits in ((92, (1 << 44) - 1), (91, 0)):
        spot[4, 6] = centre
        assert int(oracle.census_window(spot, "dense_9x5_blend")[4, 6]) == bits
    # Rows 3 above and below lie outside the dense window but inside the sparse one.
    tall = flat.copy()
    tall[4, 6] = 200
    tall[1, 6] = 0
    assert int(oracle.census_window(tall, "dense_9x5")[4, 6]) == (1 << 44) - 1


# ---- band base ------------------------------------------------------------------------------

def scipy_band_base(coarse, band, margin):
    none = np.float32(1.0e9)
    matched = np.where(coarse >= 0, coarse, none).astype(np.float32)
    lowest = ndimage.minimum_filter(matched, size=3, mode="constant", cval=none)
    start = np.clip((np.floor(lowest * 2) - margin).astype(np.int64), 0, 128 - band)
    base = np.where(lowest < none, start, -1)
    return np.repeat(np.repeat(base, 2, axis=0), 2, axis=1).astype(np.int32)


@pytest.mark.parametrize("band,margin", [(16, 4), (8, 0), (64, 7)])
def test_band_base_against_a_minimum_filter(band, margin):
    rng = np.random.default_rng(6)
    coarse = (rng.integers(0, 256, (9, 13)) / 4.0).astype(np.float32)
    coarse[rng.random((9, 13)) < 0.3] = -1.0
    assert (oracle.band_base(coarse, band, margin) == scipy_band_base(coarse, band, margin)).all()


def test_band_base_closed_forms():
    coarse = np.full((4, 5), -1.0, np.float32)
    assert (oracle.band_base(coarse) == -1).all()
    coarse[1, 2] = 10.75  # floor(21.5) - 4 = 17 over the 3x3 around it, in 2x2 blocks
    base = oracle.band_base(coarse, 16, 4)
    expected = np.full((8, 10), -1, np.int32)
    expected[0:6, 2:8] = 17
    assert (base == expected).all()
    coarse[1, 2] = 0.2  # floor(0.4) - 4 < 0: clipped to 0
    assert oracle.band_base(coarse, 16, 4)[2, 4] == 0
    coarse[1, 2] = 63.0  # 126 - 4 = 122 > 128 - 16: clipped to 112
    assert oracle.band_base(coarse, 16, 4)[2, 4] == 112


# ---- band costs, paths and selection --------------------------------------------------------

def varying_bases(rng, h, w, band, disparities):
    """Bases constant on 2x2 blocks (as coarse-to-fine makes them) with jumps of every
    size, some past a whole band, and a few single-pixel changes."""
    top = disparities - band
    blocks = rng.integers(0, top + 1, (h // 2 + 1, w // 2 + 1))
    smooth = np.repeat(np.repeat(blocks, 2, axis=0), 2, axis=1)[:h, :w]
    near = np.clip(smooth[:1, :1] + rng.integers(-3, 4, (h, w)), 0, top)
    base = np.where(rng.random((h, w)) < 0.5, smooth, near)
    lone = rng.random((h, w)) < 0.05
    base[lone] = rng.integers(0, top + 1, lone.sum())
    return base.astype(np.int32)


def test_band_costs_are_the_full_costs_at_base_plus_k():
    rng = np.random.default_rng(7)
    left = rng.integers(0, 256, (10, 40), dtype=np.uint8)
    right = rng.integers(0, 256, (10, 40), dtype=np.uint8)
    fl, fr = oracle.census(left), oracle.census(right)
    full = oracle.costs(fl, fr, 64)
    base = varying_bases(rng, 10, 40, 16, 64)
    band = oracle.band_costs(fl, fr, base, 16)
    ys, xs, ks = np.meshgrid(np.arange(10), np.arange(40), np.arange(16), indexing="ij")
    assert (band == full[ys, xs, base[:, :, None] + ks]).all()
    # Columns left of the image read a zero feature.
    assert band[0, 0, 5] == bin(int(fl[0, 0])).count("1")


@pytest.mark.parametrize("band", [8, 16])
@pytest.mark.parametrize("p1,p2", [(8, 64), (7, 70), (0, 1), (78, 79)])
def test_band_path_is_the_full_path_where_costs_outside_the_band_are_high(band, p1, p2):
    # Costs of 255 outside each pixel's band never reach inside it: inside, path values
    # stay at most p2 + 48 = 127, so the minimum is inside and an outside candidate,
    # normalized, is at least 128 > p2, worse than the jump. The full-range path then
    # holds the band path at base + k, whatever the bases do between pixels.
    rng = np.random.default_rng(band * 100 + p1)
    h, w, D = 9, 21, 40
    full = rng.integers(0, 49, (h, w, D), dtype=np.uint8)
    base = varying_bases(rng, h, w, band, D)
    ys, xs, ks = np.meshgrid(np.arange(h), np.arange(w), np.arange(band), indexing="ij")
    band_costs = full[ys, xs, base[:, :, None] + ks]
    padded = np.full_like(full, 255)
    padded[ys, xs, base[:, :, None] + ks] = band_costs
    for dy, dx in DIRECTIONS:
        expected = oracle.path(padded, dy, dx, p1, p2)[ys, xs, base[:, :, None] + ks]
        got = oracle.band_path(band_costs, base, dy, dx, p1, p2)
        assert (got == expected).all(), (dy, dx)


def test_band_aggregate_sums_the_band_paths():
    rng = np.random.default_rng(8)
    costs = rng.integers(0, 49, (8, 15, 8), dtype=np.uint8)
    base = varying_bases(rng, 8, 15, 8, 32)
    for paths, count in ((3, 3), (8, 8)):
        total = sum(oracle.band_path(costs, base, dy, dx, 8, 64) for dy, dx in DIRECTIONS[:count])
        assert (oracle.band_aggregate(costs, base, paths, 8, 64) == total).all()


def test_band_select_at_a_constant_base_is_the_shifted_full_select():
    rng = np.random.default_rng(9)
    band, D, w = 16, 64, 80
    b = D - band
    sums = rng.integers(0, 1017, (6, w, band), dtype=np.uint16)
    sums[0, 60:, 3] = 0  # a few clear winners at the band's ends and inside
    sums[1, 60:, 0] = 0
    sums[2, 60:, band - 1] = 0
    disp, keys = oracle.band_select(sums, np.full((6, w), b, np.int32), D, 1.02)
    full_disp, full_keys = oracle.select(sums[:, b:, :], 1.02)
    accepted = full_keys != 0xFFFFFFFF
    assert (disp[:, :b] == -1.0).all() and (keys[:, :b] == 0xFFFFFFFF).all()
    assert ((disp[:, b:] == -1.0) == ~accepted).all()
    # The band rounds (b + w) + c in f32 and the shifted select w + c: within an ulp at 64,
    # with the same decision to refine.
    got, shifted = disp[:, b:][accepted], full_disp[accepted]
    assert np.abs(got - (shifted + np.float32(b))).max() <= 2.0 ** -16
    assert ((got != np.floor(got)) == (shifted != np.floor(shifted))).all()
    assert (shifted != np.floor(shifted)).sum() > 50
    assert (keys[:, b:][accepted] == full_keys[accepted] + b).all()


def test_band_select_is_the_full_select_on_a_padded_volume():
    # Outside the band the full-range volume holds 0xFFFF, above every band value even
    # with the 0x8000 of d > x: the winner, second best, acceptance and key agree; the
    # subpixel step agrees inside the band and the band's ends are never refined.
    rng = np.random.default_rng(10)
    band, D, h, w = 16, 64, 7, 90
    base = varying_bases(rng, h, w, band, D)
    base[rng.random((h, w)) < 0.1] = -1
    searched = np.maximum(base, 0)
    sums = rng.integers(0, 1017, (h, w, band), dtype=np.uint16)
    edge = rng.random((h, w)) < 0.3
    sums[edge, rng.integers(0, 2, edge.sum()) * (band - 1)] = 0
    ys, xs, ks = np.meshgrid(np.arange(h), np.arange(w), np.arange(band), indexing="ij")
    padded = np.full((h, w, D), 0xFFFF, np.uint16)
    padded[ys, xs, searched[:, :, None] + ks] = sums
    disp, keys = oracle.band_select(sums, base, D, 1.02)
    full_disp, full_keys = oracle.select(padded, 1.02)
    x = np.arange(w)[None, :]
    usable = (base >= 0) & (base <= x)
    assert (disp[~usable] == -1.0).all()
    assert (keys[usable] == full_keys[usable]).all()
    accepted = usable & (keys != 0xFFFFFFFF)
    winner = (keys & 0xFFFF).astype(np.int64) - searched
    inner = accepted & (winner > 0) & (winner < band - 1)
    ends = accepted & ~inner
    assert inner.sum() > 100 and ends.sum() > 20
    assert (disp[inner] == full_disp[inner]).all()
    assert (disp[ends] == (keys[ends] & 0xFFFF).astype(np.float32)).all()


# ---- sgm_band and coarse-to-fine (nx_stereo_test.exs:115-225) -------------------------------

@pytest.mark.parametrize("paths", [3, 8])
def test_a_band_covering_every_disparity_matches_as_sgm(paths):
    left, right = layered(24, 160, 4.5, 14.25)
    opts = dict(p1=7, p2=70, uniqueness=1.05, paths=paths)
    band = oracle.sgm_band(left, right, np.zeros((24, 160), np.int32), 64, 64, **opts)
    assert (band == oracle.sgm_raw(left, right, 64, **opts)).all()


def test_a_band_around_each_pixels_disparity_finds_the_full_ranges_matches():
    left, right = layered(24, 160, 4.5, 14.25)
    opts = dict(p1=7, p2=70, uniqueness=1.05)
    full = oracle.sgm_raw(left, right, 64, **opts)
    base = np.clip(np.floor(np.maximum(full, 0)).astype(np.int32) - 4, 0, 48)
    band = oracle.sgm_band(left, right, base, 64, 16, **opts)
    both = (band >= 0) & (full >= 0)
    assert both.sum() >= 0.9 * (full >= 0).sum()
    assert (np.abs(band - full)[both] <= 0.5).sum() >= 0.99 * both.sum()


def test_a_negative_base_gives_no_match():
    left, right = pair(16, 80, 3)
    base = np.zeros((16, 80), np.int32)
    base[:, 40:60] = -1
    band = oracle.sgm_band(left, right, base, 64, 16, p1=10, p2=60)
    assert (band[:, 40:60] == -1.0).all()
    assert (band[4:12, 10:31] > 0).all()


def test_coarse_to_fine_matches_a_shifted_texture_as_the_full_search():
    left, right = pair(32, 240, 37)
    fine = oracle.coarse_to_fine(left, right)
    full = oracle.sgm_raw(left, right, 128)
    assert (interior(fine, 140) == interior(full, 140)).all()
    assert (np.abs(interior(fine, 140) - 37.0) < 0.1).sum() >= 2300


def test_coarse_to_fine_agrees_with_the_full_search_across_depth_edges():
    left, right = layered(48, 240, 20.5, 60.25)
    opts = dict(p1=7, p2=70, uniqueness=1.05)
    full = oracle.sgm_raw(left, right, 128, **opts)
    fine = oracle.coarse_to_fine(left, right, **opts)
    both = (fine >= 0) & (full >= 0)
    assert both.sum() >= 0.9 * (full >= 0).sum()
    assert (np.abs(fine - full)[both] <= 1).sum() >= 0.95 * both.sum()


@pytest.mark.parametrize("census", ["dense_9x5", "dense_9x5_blend"])
def test_coarse_to_fine_with_dense_census_finds_the_shift(census):
    left, right = pair(32, 240, 37)
    fine = oracle.coarse_to_fine(left, right, fine_census=census, band=32, margin=8,
                                 coarse_paths=3)
    assert (np.abs(interior(fine, 140) - 37.0) < 0.1).mean() >= 0.98


def test_coarse_to_fine_is_the_band_search_from_the_half_size_match():
    # nx_stereo.ex:186-213: shrink, the 64-disparity match of the halves with sgm/3's own
    # output (a match at 0 kept), band_base, then sgm_band at 128 disparities.
    left, right = pair(32, 240, 37)
    coarse = oracle.sgm_raw(oracle.shrink(left), oracle.shrink(right), 64)
    base = oracle.band_base(coarse, 64, 63)
    assert base.min() >= -1 and base.max() <= 64
    fine = oracle.coarse_to_fine(left, right, band=64, margin=63)
    assert (fine == oracle.sgm_band(left, right, base, 128, 64)).all()


# ---- resize_area (Decode, image.ex:402-476) -------------------------------------------------

def test_resize_area_twofold_is_the_rounded_block_mean():
    # Twofold: weights 0.5 and 0.5, every sum a quarter of an integer, exact in f32; round
    # half away from zero is then (sum + 2) // 4.
    frame = np.random.default_rng(11).integers(0, 256, (18, 26), dtype=np.uint8)
    assert (oracle.resize_area(frame, (9, 13)) == oracle.shrink(frame)).all()


def test_resize_area_closed_forms():
    frame = np.random.default_rng(12).integers(0, 256, (7, 9), dtype=np.uint8)
    assert (oracle.resize_area(frame, (7, 9)) == frame).all()
    assert (oracle.resize_area(np.full((21, 33), 201, np.uint8), (8, 10)) == 201).all()
    # 3 -> 2 columns: taps (0, 2/3), (1, 1/3) and (1, 1/3), (2, 2/3).
    row = np.array([[30, 60, 90]], np.uint8)
    assert oracle.resize_area(row, (1, 2)).tolist() == [[40, 80]]
    # 5 -> 2 rows: taps (0, .4) (1, .4) (2, .2) and (2, .2) (3, .4) (4, .4).
    column = np.array([[10], [20], [100], [40], [50]], np.uint8)
    assert oracle.resize_area(column, (2, 1)).tolist() == [[32], [56]]
    # 4 -> 3 of a ramp: taps (0, .75) (1, .25); (1, .5) (2, .5); (2, .25) (3, .75).
    ramp = np.array([[0, 40, 80, 120]], np.uint8)
    assert oracle.resize_area(ramp, (1, 3)).tolist() == [[10, 60, 110]]
