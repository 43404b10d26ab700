"""Synthetic pairs of the reference's matcher tests (test/nx_stereo_test.exs:5-56), drawn
from numpy's generator instead of Erlang's :rand."""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def pair(height, width, shift, seed=1):
    """nx_stereo_test.exs:5-12: a random texture and its view `shift` columns on."""
    wide = np.random.default_rng(seed).integers(0, 256, (height, width + shift), dtype=np.uint8)
    return wide[:, :width].copy(), wide[:, shift:shift + width].copy()


def _smooth_texture(rng, height, width):
    """nx_stereo_test.exs:34-42: uniform noise under the 3x3 binomial kernel, zero padded."""
    noise = rng.random((height, width)) * 255.0
    kernel = np.array([[1.0, 2.0, 1.0], [2.0, 4.0, 2.0], [1.0, 2.0, 1.0]]) / 16
    return ndimage.correlate(noise, kernel, mode="constant", cval=0.0)


def _view(texture, width, shift):
    """nx_stereo_test.exs:44-51: texture[y, x + shift], linearly interpolated."""
    i = int(shift)
    f = shift - i
    return texture[:, i:i + width] * (1 - f) + texture[:, i + 1:i + 1 + width] * f


def layered(height, width, back, front, seed=1):
    """nx_stereo_test.exs:14-32: a smooth background at disparity `back` behind a band of
    columns at `front` (fractional disparities next to occlusions)."""
    rng = np.random.default_rng(seed)
    background = _smooth_texture(rng, height, width + 40)
    band = _smooth_texture(rng, height, width // 3 + 40)
    x, n = width // 3, width // 3
    left = _view(background, width, 0)
    left[:, x:x + n] = _view(band, n, 0)
    right = _view(background, width, back)
    at = x - int(front)
    right[:, at:at + n] = _view(band, n, front - int(front))
    # Nx.round: halves away from zero (the values are non-negative)
    return (np.floor(left + 0.5).astype(np.uint8), np.floor(right + 0.5).astype(np.uint8))


def interior(disparity, margin_x):
    """nx_stereo_test.exs:53-56."""
    height, width = disparity.shape
    return disparity[4:height - 4, margin_x:width - 4].ravel()
