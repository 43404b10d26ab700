"""Band search and coarse-to-fine on the GPU (nxs_match_band_host, nxs_coarse_to_fine_host)
against the oracle, stage by stage and bit for bit (nx_stereo.ex:164-471)."""
from __future__ import annotations

import numpy as np
import pytest

from nx_stereo_amd import _native, scenes
from oracle import oracle
from tests.band_scenes import interior, layered, pair
from tests.test_oracle_band import varying_bases

pytestmark = pytest.mark.gpu


def oracle_band_stages(left, right, base, disparities, band, paths, census, p1, p2, uniqueness):
    fl = oracle.census_window(left, census)
    fr = oracle.census_window(right, census)
    searched = np.maximum(base, 0)
    costs = oracle.band_costs(fl, fr, searched, band)
    sums = oracle.band_aggregate(costs, searched, paths, p1, p2)
    selected, keys = oracle.band_select(sums, base, disparities, uniqueness)
    checked = oracle.consistency(selected, keys, disparities)
    return dict(census_left=fl, census_right=fr, costs=costs, sums=sums, selected=selected,
                keys=keys, checked=checked, disparity=oracle.median3(checked))


def assert_stages_equal(got, expected):
    for name, want in expected.items():
        have = got[name]
        assert have.shape == want.shape, name
        bad = np.argwhere(have != want) if have.dtype != np.float32 else \
            np.argwhere(have.view(np.uint32) != want.view(np.uint32))
        assert len(bad) == 0, f"{name}: {len(bad)} differ, first at {bad[:3].tolist()}"


@pytest.mark.parametrize("paths", [3, 8])
@pytest.mark.parametrize("band", [8, 16, 32, 64])
def test_band_stages_match_the_oracle(band, paths):
    print(f"band {band} paths {paths}", flush=True)
    left, right = layered(40, 200, 6.5, 30.25, seed=band + paths)
    disparities = 128 if band == 64 else 64
    rng = np.random.default_rng(band * 10 + paths)
    base = varying_bases(rng, 40, 200, band, disparities)
    base[rng.random(base.shape) < 0.05] = -1
    base[:6, 150:170] = -1
    opts = dict(p1=7, p2=70, uniqueness=1.05)
    got = _native.sgm_band(left, right, base, disparities, band, paths, "sparse_9x7",
                           intermediates=True, **opts)
    expected = oracle_band_stages(left, right, base, disparities, band, paths, "sparse_9x7",
                                  **opts)
    assert_stages_equal(got, expected)
    assert (got["disparity"] > 0).mean() > 0.2


@pytest.mark.parametrize("census", ["dense_9x5", "dense_9x5_blend"])
def test_band_census_windows_match_the_oracle(census):
    print(f"census {census}", flush=True)
    left, right, _ = scenes.planar_pair(160, 48, 7, low=4.0, high=40.0)
    base = varying_bases(np.random.default_rng(3), 48, 160, 16, 64)
    got = _native.sgm_band(left, right, base, 64, 16, 8, census, intermediates=True)
    expected = oracle_band_stages(left, right, base, 64, 16, 8, census, 8, 64,
                                  1.0526315789473684)
    assert_stages_equal(got, expected)


def test_a_band_covering_every_disparity_is_the_full_search():
    left, right = layered(24, 160, 4.5, 14.25)
    opts = dict(p1=7, p2=70, uniqueness=1.05)
    got = _native.sgm_band(left, right, np.zeros((24, 160), np.int32), 64, 64, 8, **opts)
    assert (got["disparity"] == oracle.sgm_raw(left, right, 64, 8, **opts)).all()


def test_bases_far_past_the_image_match_nothing():
    left, right = pair(16, 80, 3)
    base = np.zeros((16, 80), np.int32)
    base[:, 30:50] = 60000  # every candidate left of the image
    base[:, 50:60] = 2 ** 31 - 1
    got = _native.sgm_band(left, right, base, 64, 16, 8, p1=10, p2=60)["disparity"]
    assert (got[:, 31:59] == -1.0).all()
    expected = oracle.sgm_band(left, right, np.minimum(base, 1 << 16), 64, 16, 8, p1=10, p2=60)
    assert (got == expected).all()


C2F_CASES = [
    ((32, 240), {}),
    ((48, 240), dict(p1=7, p2=70, uniqueness=1.05)),
    ((64, 200), dict(band=32, margin=8, coarse_paths=3)),
    ((56, 180), dict(band=8, margin=2, paths=3, fine_census="dense_9x5_blend")),
    ((40, 160), dict(band=64, margin=0, fine_census="dense_9x5", p1=0, p2=1)),
    ((544, 728), {}),
    ((544, 728), dict(paths=3, coarse_paths=3, band=32)),
    ((1080, 1920), {}),
]


@pytest.mark.parametrize("size,opts", C2F_CASES)
def test_coarse_to_fine_matches_the_oracle(size, opts):
    h, w = size
    print(f"coarse_to_fine {w}x{h} {opts}", flush=True)
    left, right, _ = scenes.planar_pair(w, h, 11, low=4.0, high=100.0)
    got = _native.sgm_coarse_to_fine(left, right, intermediates=True, **opts)
    band, margin = opts.get("band", 16), opts.get("margin", 4)
    paths = opts.get("paths", 8)
    coarse_paths = opts.get("coarse_paths", paths)
    penalties = {k: opts[k] for k in ("p1", "p2", "uniqueness") if k in opts}
    half_left, half_right = oracle.shrink(left), oracle.shrink(right)
    assert (got["half_left"] == half_left).all() and (got["half_right"] == half_right).all()
    coarse = oracle.sgm_raw(half_left, half_right, 64, coarse_paths, **penalties)
    assert (got["coarse"].view(np.uint32) == coarse.view(np.uint32)).all()
    assert (got["base"] == oracle.band_base(coarse, band, margin)).all()
    expected = oracle.coarse_to_fine(left, right, **opts)
    assert (got["disparity"].view(np.uint32) == expected.view(np.uint32)).all()


def test_coarse_to_fine_finds_a_shifted_texture():
    # nx_stereo_test.exs:116-125, on numpy's texture.
    left, right = pair(32, 240, 37)
    fine = _native.sgm_coarse_to_fine(left, right)["disparity"]
    full = oracle.sgm_raw(left, right, 128)
    assert (interior(fine, 140) == interior(full, 140)).all()
    assert (np.abs(interior(fine, 140) - 37.0) < 0.1).sum() >= 2300


def test_the_library_refuses_what_sgm_band_and_coarse_to_fine_refuse():
    left, right = pair(16, 80, 3)
    base = np.zeros((16, 80), np.int32)
    for kwargs, reason in ((dict(disparities=192), "disparities must be 64 or 128"),
                           (dict(paths=4), "paths must be 3 or 8"),
                           (dict(p1=10, p2=10), "0 <= p1 < p2")):
        with pytest.raises(_native.NativeError, match=reason):
            _native.sgm_band(left, right, base, **kwargs)
    for kwargs, reason in ((dict(margin=-1), "margin must be"),
                           (dict(coarse_paths=5), "paths must be 3 or 8")):
        with pytest.raises(_native.NativeError, match=reason):
            _native.sgm_coarse_to_fine(*pair(32, 80, 3), **kwargs)
    # Halves of 16 x 8 are too small for the 64-disparity match (wider than 8 needed).
    with pytest.raises(_native.NativeError, match="wider than 8"):
        _native.sgm_coarse_to_fine(*pair(16, 16, 3))


EXTREME_CASES = [
    ((14, 4096), dict(band=64, margin=3)),            # the widest image, the lowest height
    ((14, 4096), dict(band=8, paths=3, coarse_paths=3)),
    ((14, 18), {}),                                   # the smallest: halves of 9 x 7
    ((30, 250), dict(band=32, fine_census="dense_9x5")),  # widths off every line multiple
    ((202, 66), dict(band=16, margin=0)),             # taller than wide
]


@pytest.mark.parametrize("size,opts", EXTREME_CASES)
def test_coarse_to_fine_at_the_extremes(size, opts):
    h, w = size
    print(f"coarse_to_fine extreme {w}x{h} {opts}", flush=True)
    left, right, _ = scenes.planar_pair(w, h, 13, low=1.0, high=min(110.0, w - 6))
    got = _native.sgm_coarse_to_fine(left, right, intermediates=True, **opts)
    expected = oracle.coarse_to_fine(left, right, **opts)
    assert (got["disparity"].view(np.uint32) == expected.view(np.uint32)).all()


@pytest.mark.parametrize("band", [16, 64])
def test_band_search_at_the_widest_image(band):
    h, w = 9, 4096
    print(f"band {band} at {w}x{h}", flush=True)
    left, right, _ = scenes.planar_pair(w, h, 14, low=2.0, high=100.0)
    base = varying_bases(np.random.default_rng(band), h, w, band, 128)
    got = _native.sgm_band(left, right, base, 128, band, 8, intermediates=True)
    expected = oracle_band_stages(left, right, base, 128, band, 8, "sparse_9x7", 8, 64,
                                  1.0526315789473684)
    assert_stages_equal(got, expected)
