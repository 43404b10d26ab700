"""Seeded synthetic stereo inputs.

* ``noise`` and ``scene``: the reference's own reproducible test images,
  restated bit for bit (native/stereo_match/src/lib.rs:448-508): a xorshift
  generator, a smooth texture seen through a slanted plane, and a flat patch.
* ``planar_pair``: rectified pairs of a piecewise-planar scene. The left image
  is seeded uniform noise under a 3x3 box blur; the right one is the left
  warped by a disparity field of 4-8 fronto-parallel or slanted planes over a
  Voronoi partition, plus Gaussian noise, rounded and clamped to u8.
* ``rig`` and ``raw_frames``: a calibrated camera pair (f, principal point at
  the image centre, plumb_bob k1 and k2, rectifying rotations half a degree
  apart, a baseline) and unrectified frames whose rectification is a planar
  pair.
"""
from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


class Random:
    """lib.rs:448-465: xorshift on u64, seeded through a golden-ratio multiply."""

    def __init__(self, seed: int):
        self.state = ((seed * 0x9E3779B97F4A7C15) & _MASK) | 1

    def next(self) -> int:
        x = self.state
        x ^= (x << 13) & _MASK
        x ^= x >> 7
        x ^= (x << 17) & _MASK
        self.state = x
        return x

    def byte(self) -> int:
        return (self.next() >> 24) & 0xFF

    def bytes(self, count: int) -> np.ndarray:
        return np.fromiter((self.byte() for _ in range(count)), np.uint8, count)


def noise(width: int, height: int, seed: int) -> np.ndarray:
    """lib.rs:467-470, as an {height, width} u8 image."""
    return Random(seed).bytes(width * height).reshape(height, width)


def scene(width: int, height: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """lib.rs:475-508: a slanted-plane texture with a flat patch, in f32 as the reference."""
    f32 = np.float32
    random = Random(seed)
    texture_width = width + 160
    coarse = random.bytes(texture_width * height).astype(f32).reshape(height, texture_width)

    def texture(y, x):
        x0 = np.clip(np.floor(x), f32(0), f32(texture_width - 2)).astype(np.int64)
        weight = (x - x0.astype(f32)).astype(f32)
        rows = np.arange(height)[:, None]
        a = coarse[rows, x0] * (f32(1) - weight)
        b = coarse[rows, x0 + 1] * weight
        return (a + b).astype(f32)

    xs = np.arange(width, dtype=f32)[None, :].repeat(height, 0)
    ys = np.arange(height, dtype=f32)[:, None].repeat(width, 1)
    disparity = (f32(4.0) + (f32(30.0) * xs) / f32(width)) + (f32(6.0) * ys) / f32(height)
    xi, yi = np.meshgrid(np.arange(width), np.arange(height))
    flat = (xi > width // 3) & (xi < width // 3 + 10) & (yi > height // 2) & (yi < height // 2 + 8)
    left_x = (xs + f32(100.0)).astype(f32)
    right_x = ((xs + f32(100.0)) + disparity).astype(f32)
    left = np.where(flat, 90, texture(None, left_x).astype(np.uint8)).astype(np.uint8)
    right = np.where(flat, 90, texture(None, right_x).astype(np.uint8)).astype(np.uint8)
    return left, right


def _box3(image: np.ndarray) -> np.ndarray:
    padded = np.pad(image.astype(np.int32), 1, mode="edge")
    h, w = image.shape
    total = sum(padded[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3))
    return ((total + 4) // 9).astype(np.uint8)


def disparity_field(width: int, height: int, rng: np.random.Generator, low: float, high: float,
                    planes: tuple[int, int] = (4, 8)) -> np.ndarray:
    """A piecewise-planar f32 field: Voronoi cells, each fronto-parallel or slanted."""
    count = int(rng.integers(planes[0], planes[1] + 1))
    seeds = np.stack([rng.uniform(0, width, count), rng.uniform(0, height, count)], 1)
    ys, xs = np.mgrid[0:height, 0:width].astype(np.float64)
    distance = (xs[None] - seeds[:, 0, None, None]) ** 2 + (ys[None] - seeds[:, 1, None, None]) ** 2
    cell = np.argmin(distance, axis=0)
    field = np.empty((height, width), np.float64)
    span = high - low
    for k in range(count):
        base = rng.uniform(low + 0.1 * span, high - 0.1 * span)
        if rng.random() < 0.5:
            gx = gy = 0.0
        else:
            gx = rng.uniform(-0.1, 0.1) * span / width
            gy = rng.uniform(-0.1, 0.1) * span / height
        cx, cy = seeds[k]
        field[cell == k] = (base + gx * (xs - cx) + gy * (ys - cy))[cell == k]
    return np.clip(field, low, high).astype(np.float32)


def _bilinear(image: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
    h, w = image.shape
    x0 = np.floor(xs).astype(np.int64)
    y0 = np.floor(ys).astype(np.int64)
    fx, fy = xs - x0, ys - y0
    x0c, x1c = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    y0c, y1c = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    img = image.astype(np.float64)
    top = img[y0c, x0c] * (1 - fx) + img[y0c, x1c] * fx
    bottom = img[y1c, x0c] * (1 - fx) + img[y1c, x1c] * fx
    return top * (1 - fy) + bottom * fy


def planar_pair(width: int, height: int, seed: int, low: float = 4.0, high: float = 60.0,
                sigma: float = 2.0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """A rectified (left, right, true disparity) triple from the planar-scene generator.

    right[y, xr] samples left at xr + d(y, xr): a point at column x in the left image
    lies at x - d in the right one.
    """
    rng = np.random.default_rng(seed)
    # The texture is wider than the image so the right view has content past the left edge.
    margin = int(math.ceil(high)) + 2
    base = rng.integers(0, 256, size=(height, width + margin), dtype=np.uint8)
    texture = _box3(base)
    left = texture[:, margin:]
    field = disparity_field(width, height, rng, low, high)
    ys, xs = np.mgrid[0:height, 0:width].astype(np.float64)
    sampled = _bilinear(texture, xs + field + margin, ys)
    right = sampled + rng.normal(0.0, sigma, size=sampled.shape)
    right = np.clip(np.round(right), 0, 255).astype(np.uint8)
    return np.ascontiguousarray(left), right, field


# --------------------------------------------------------------------------------------------
# A calibrated rig and the raw frames that rectify to a planar pair.


def _rotation_y(degrees: float) -> np.ndarray:
    t = math.radians(degrees)
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def rig(width: int, height: int, focal: float = 873.0, baseline: float = 0.2,
        k1: float = -0.05, k2: float = 0.01, angle_deg: float = 0.5) -> dict:
    """A stereo rig as a calibration tool reports it: K, plumb_bob d, R and P per camera.

    The rectifying rotations are +-angle/2 about the vertical axis; the rectified camera
    matrix is K's, and the right projection carries -f * baseline in P[0][3].
    """
    cx, cy = width / 2.0, height / 2.0
    k = [focal, 0.0, cx, 0.0, focal, cy, 0.0, 0.0, 1.0]
    d = [k1, k2, 0.0, 0.0, 0.0]
    p_left = [focal, 0.0, cx, 0.0, 0.0, focal, cy, 0.0, 0.0, 0.0, 1.0, 0.0]
    p_right = [focal, 0.0, cx, -focal * baseline, 0.0, focal, cy, 0.0, 0.0, 0.0, 1.0, 0.0]
    r_left = _rotation_y(angle_deg / 2).reshape(-1).tolist()
    r_right = _rotation_y(-angle_deg / 2).reshape(-1).tolist()
    return {
        "width": width, "height": height, "focal_length_px": focal, "baseline_m": baseline,
        "principal_point": (cx, cy),
        "left": {"k": k, "d": d, "r": r_left, "p": p_left},
        "right": {"k": k, "d": d, "r": r_right, "p": p_right},
    }


def camera(k, d, r, p) -> dict:
    """rectify.ex:34-50: k, d and inverse(P3 R) computed in f64 then cast to f32."""
    p3 = np.array([p[i] for i in (0, 1, 2, 4, 5, 6, 8, 9, 10)], np.float64).reshape(3, 3)
    pr = p3 @ np.array(r, np.float64).reshape(3, 3)
    a, b, c, dd, e, f, g, h, i = pr.reshape(-1)
    cof = [e * i - f * h, c * h - b * i, b * f - c * e, f * g - dd * i, a * i - c * g,
           c * dd - a * f, dd * h - e * g, b * g - a * h, a * e - b * dd]
    det = a * cof[0] + b * cof[3] + c * cof[6]
    inverse = np.array([v / det for v in cof], np.float64)
    coefficients = (list(d) + [0.0] * 5)[:5]
    return {"k": np.array(k, np.float32), "d": np.array(coefficients, np.float32),
            "inverse_pr": inverse.astype(np.float32)}


def raw_frames(calibration: dict, seed: int, low: float | None = None,
               high: float = 127.0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Unrectified (left_frame, right_frame) of a planar scene, and its rectified disparity.

    The scene is a planar pair in rectified coordinates (disparities f*B/Z for planes
    at 0.5-15 m, clipped to what `high` disparities can match); each frame pixel
    takes the rectified image at the point the camera model maps it to.
    """
    width, height = calibration["width"], calibration["height"]
    f, b = calibration["focal_length_px"], calibration["baseline_m"]
    if low is None:
        low = max(4.0, f * b / 15.0)
    left, right, field = planar_pair(width, height, seed, low=low, high=high)
    frames = []
    for side, image in (("left", left), ("right", right)):
        cam = calibration[side]
        k = np.array(cam["k"], np.float64).reshape(3, 3)
        dist = np.array(cam["d"], np.float64)
        r = np.array(cam["r"], np.float64).reshape(3, 3)
        p3 = np.array([cam["p"][i] for i in (0, 1, 2, 4, 5, 6, 8, 9, 10)]).reshape(3, 3)
        vs, us = np.mgrid[0:height, 0:width].astype(np.float64)
        # Frame pixel -> distorted normalised -> undistorted (fixed point) -> rectified pixel.
        xd = (us - k[0, 2]) / k[0, 0]
        yd = (vs - k[1, 2]) / k[1, 1]
        x, y = xd.copy(), yd.copy()
        for _ in range(20):
            r2 = x * x + y * y
            radial = 1 + dist[0] * r2 + dist[1] * r2 * r2 + dist[4] * r2 ** 3
            dx = 2 * dist[2] * x * y + dist[3] * (r2 + 2 * x * x)
            dy = dist[2] * (r2 + 2 * y * y) + 2 * dist[3] * x * y
            x = (xd - dx) / radial
            y = (yd - dy) / radial
        # The rectified pixel of camera ray q is P3 R q (rectify.ex maps it back with inverse(P3 R)).
        rays = np.stack([x, y, np.ones_like(x)], -1) @ r.T
        projected = rays @ p3.T
        ur = projected[..., 0] / projected[..., 2]
        vr = projected[..., 1] / projected[..., 2]
        sampled = _bilinear(image, ur, vr)
        outside = (ur < -0.5) | (ur > width - 0.5) | (vr < -0.5) | (vr > height - 0.5)
        frame = np.where(outside, 0, np.clip(np.round(sampled), 0, 255)).astype(np.uint8)
        frames.append(frame)
    return frames[0], frames[1], field
