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
