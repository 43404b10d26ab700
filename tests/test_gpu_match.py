"""The matcher on the GPU against the oracle, bit for bit: census features,
matching costs, aggregated path costs, selection (disparity and key), the
left-right check and the stage's output, on the reference's own test scenes
(native/stereo_match/src/lib.rs:535-568) and on planar-scene pairs."""
import numpy as np
import pytest

from nx_stereo_amd import _native, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{len(diff)} differ, first at {tuple(diff[0])}: {a[tuple(diff[0])]} vs {b[tuple(diff[0])]}"


def check_all(left, right, disparities, paths, p1=8, p2=64, uniqueness=1.0526315789473684):
    print(f"case {left.shape} D={disparities} paths={paths} p1={p1} p2={p2} u={uniqueness}", flush=True)
    got = _native.match_host(left, right, disparities, paths, p1, p2, uniqueness, intermediates=True)
    fl, fr = oracle.census(left), oracle.census(right)
    same_bits(got["census_left"], fl)
    same_bits(got["census_right"], fr)
    costs = oracle.costs(fl, fr, disparities)
    same_bits(got["costs"], costs)
    sums = oracle.aggregate(costs, paths, p1, p2)
    same_bits(got["sums"], sums)
    sel, keys = oracle.select(sums, uniqueness)
    same_bits(got["selected"], sel)
    same_bits(got["keys"], keys)
    checked = oracle.consistency(sel, keys, disparities)
    same_bits(got["checked"], checked)
    want = oracle.sgm(left, right, disparities, paths, p1, p2, uniqueness)
    same_bits(got["disparity"], want)


@pytest.mark.parametrize("size", [(64, 48), (97, 31)])
@pytest.mark.parametrize("disparities", [64, 128])
@pytest.mark.parametrize("paths", [3, 8])
def test_reference_scenes_bit_exact(size, disparities, paths):
    w, h = size
    check_all(*scenes.scene(w, h, 7), disparities, paths)
    check_all(scenes.noise(w, h, 1), scenes.noise(w, h, 2), disparities, paths)


@pytest.mark.parametrize("p1,p2,uniqueness", [(0, 1, 1.0), (4, 40, 1.2), (78, 79, 2.5), (10, 79, 1.0526316)])
def test_other_penalties_and_uniqueness(p1, p2, uniqueness):
    left, right = scenes.scene(80, 40, 3)
    for paths in (3, 8):
        check_all(left, right, 64, paths, p1, p2, uniqueness)


@pytest.mark.parametrize("paths", [3, 8])
@pytest.mark.parametrize("disparities", [192, 256])
def test_wide_disparity_ranges(disparities, paths):
    left, right, _ = scenes.planar_pair(300, 40, seed=5, low=4, high=disparities - 10)
    check_all(left, right, disparities, paths)


@pytest.mark.parametrize("size", [(9, 7), (10, 8), (33, 17), (17, 40)])
def test_small_and_odd_sizes(size):
    """Lines of every length from 1 up, odd and even, meet in the middle correctly."""
    w, h = size
    for paths in (3, 8):
        check_all(scenes.noise(w, h, 3), scenes.noise(w, h, 4), 64, paths)


def test_planar_pair_364x272x64():
    left, right, _ = scenes.planar_pair(364, 272, seed=1)
    check_all(left, right, 64, 8)


def test_maximum_width():
    left, right, _ = scenes.planar_pair(4096, 8, seed=9, low=4, high=50)
    got = _native.match_host(left, right, 64, 8)["disparity"]
    same_bits(got, oracle.sgm(left, right, 64, 8))


@pytest.mark.parametrize("size,disparities", [((20, 600), 64), ((37, 530), 128), ((13, 1100), 192),
                                              ((70, 520), 256)])
def test_tall_narrow_images(size, disparities):
    """Five hundred rows and more of narrow images (1024-thread census blocks, widths off a
    multiple of four, fewer pixels than disparities): census, costs and the disparity bit for
    bit."""
    w, h = size
    left, right = scenes.scene(w, h, 11)
    got = _native.match_host(left, right, disparities, 8, intermediates=True)
    fl, fr = oracle.census(left), oracle.census(right)
    same_bits(got["census_left"], fl)
    same_bits(got["census_right"], fr)
    same_bits(got["costs"], oracle.costs(fl, fr, disparities))
    same_bits(got["disparity"], oracle.sgm(left, right, disparities, 8))
