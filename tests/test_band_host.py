"""The host mirror of sgm_band/4 and sgm_coarse_to_fine/3 refuses what the reference refuses
(nx_stereo_test.exs:139-147, 208-225), before any device work."""
from __future__ import annotations

import numpy as np
import pytest

from nx_stereo_amd import _native
from tests.band_scenes import pair


def test_sgm_band_validates_the_band_and_the_base():
    left, right = pair(16, 80, 3)
    base = np.zeros((16, 80), np.int32)
    for band in (12, 128):
        with pytest.raises(ValueError, match=r":band must be 8, 16, 32 or 64"):
            _native.sgm_band(left, right, base, band=band)
    with pytest.raises(ValueError, match=r"base.*s32.*\(16, 80\)"):
        _native.sgm_band(left, right, np.zeros((16, 79), np.int32), band=16)
    with pytest.raises(ValueError, match=r"base.*s32"):
        _native.sgm_band(left, right, base.astype(np.int64), band=16)
    with pytest.raises(ValueError, match=r":census must be one of"):
        _native.sgm_band(left, right, base, census="dense_7x7")


def test_sgm_coarse_to_fine_validates_its_inputs():
    left, right = pair(16, 80, 3)
    with pytest.raises(ValueError, match="even height and width"):
        _native.sgm_coarse_to_fine(left[:15], right[:15])
    with pytest.raises(ValueError, match=":disparities must be 128"):
        _native.sgm_coarse_to_fine(left, right, disparities=64)
    with pytest.raises(ValueError, match=":fine_census must be one of"):
        _native.sgm_coarse_to_fine(left, right, fine_census="sparse")


def test_the_disparity_stage_checks_coarse_to_fine_as_its_init_does():
    from nx_stereo_amd.stages import Disparity, StageError

    with pytest.raises(StageError, match="coarse_to_fine searches 128 disparities, got 64"):
        Disparity(search="coarse_to_fine")
    with pytest.raises(StageError, match=r"margin must be below band \(8\), got 8"):
        Disparity(search="coarse_to_fine", num_disparities=128, band=8, margin=8)
    with pytest.raises(StageError, match="fine_census requires coarse_to_fine search"):
        Disparity(fine_census="dense_9x5")
    with pytest.raises(StageError, match="band must be one of"):
        Disparity(search="coarse_to_fine", num_disparities=128, band=12)
    stage = Disparity(search="coarse_to_fine", num_disparities=128, paths=3,
                      fine_census="dense_9x5_blend")
    assert stage.state["coarse_paths"] == 3  # paths by default
    assert stage.native_settings() == [
        ("num_disparities", 128), ("paths", 3), ("search", 1), ("coarse_paths", 3),
        ("band", 16), ("margin", 4), ("fine_census", 2)]
    assert Disparity(num_disparities=128).native_settings() == [
        ("num_disparities", 128), ("paths", 8)]
    assert stage.compute() is None
