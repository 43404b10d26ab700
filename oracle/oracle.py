"""ctypes bindings of the oracle and of the CPU baseline.

Test and benchmark infrastructure only: the tests, the smoke check and the
benchmark's CPU-baseline leg load these libraries as the checker. The
product package never imports this module.

The C sources (oracle.c, cpu_baseline.c) cite the reference file:line each
function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p


def build() -> None:
    """Compiles liboracle.so and libcpubaseline.so beside their sources."""
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C oracle`")
    return ctypes.CDLL(path)


_lib = None
_cpu = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = _load("liboracle.so")
        i, f = ctypes.c_int, ctypes.c_float
        _lib.orc_census.argtypes = [_P, i, i, _P]
        _lib.orc_costs.argtypes = [_P, _P, i, i, i, _P]
        _lib.orc_path.argtypes = [_P, i, i, i, i, i, i, i, _P]
        _lib.orc_aggregate.argtypes = [_P, i, i, i, i, i, i, _P]
        _lib.orc_select.argtypes = [_P, i, i, i, f, _P, _P]
        _lib.orc_consistency.argtypes = [_P, _P, i, i, i, _P]
        _lib.orc_median3.argtypes = [_P, i, i, _P]
        _lib.orc_sgm.argtypes = [_P, _P, i, i, i, i, i, i, f, _P]
        _lib.orc_sgm.restype = i
        _lib.orc_rectify.argtypes = [_P, i, i, _P, _P, _P, i, i, _P]
        _lib.orc_despeckle.argtypes = [_P, i, i, f, f, i, _P]
        _lib.orc_depth.argtypes = [_P, i, f, f, _P, _P]
        _lib.orc_cloud.argtypes = [_P, i, i, i, f, f, f, f, f, _P]
        _lib.orc_cloud.restype = i
        _lib.orc_temporal.argtypes = [_P, _P, _P, i, f, f, f, f]
        _lib.orc_quality.argtypes = [_P, _P, i, i, _P]
        _lib.orc_median.argtypes = [_P, i, i, i, _P]
        _lib.orc_sharpness.argtypes = [_P, _P, i, i, _P]
        _lib.orc_alignment.argtypes = [_P, _P, _P, i, i, i, i, i, _P]
        _lib.orc_resize_area.argtypes = [_P, i, i, i, i, _P]
        _lib.orc_sgm_raw.argtypes = [_P, _P, i, i, i, i, i, i, f, _P]
        _lib.orc_sgm_raw.restype = i
        _lib.orc_shrink.argtypes = [_P, i, i, _P]
        _lib.orc_census_window.argtypes = [_P, i, i, i, _P]
        _lib.orc_band_base.argtypes = [_P, i, i, i, i, _P]
        _lib.orc_band_costs.argtypes = [_P, _P, _P, i, i, i, _P]
        _lib.orc_band_path.argtypes = [_P, _P, i, i, i, i, i, i, i, _P]
        _lib.orc_band_aggregate.argtypes = [_P, _P, i, i, i, i, i, i, _P]
        _lib.orc_band_select.argtypes = [_P, _P, i, i, i, i, f, _P, _P]
        _lib.orc_sgm_band.argtypes = [_P, _P, _P, i, i, i, i, i, i, i, i, f, _P]
        _lib.orc_sgm_band.restype = i
        _lib.orc_coarse_to_fine.argtypes = [_P, _P, i, i, i, i, i, i, i, i, i, f, i, _P]
        _lib.orc_coarse_to_fine.restype = i
    return _lib


def cpu() -> ctypes.CDLL:
    global _cpu
    if _cpu is None:
        _cpu = _load("libcpubaseline.so")
        i, f = ctypes.c_int, ctypes.c_float
        _cpu.cpu_sgm.argtypes = [_P, _P, i, i, i, i, i, i, f, _P, i]
        _cpu.cpu_sgm.restype = i
        _cpu.cpu_rectify_pair.argtypes = [_P, _P, i, i] + [_P] * 6 + [i, i, _P, _P, i]
        _cpu.cpu_rectify_pair.restype = i
        _cpu.cpu_despeckle.argtypes = [_P, i, i, f, f, i, _P, i]
        _cpu.cpu_despeckle.restype = i
        _cpu.cpu_depth.argtypes = [_P, i, i, f, f, _P, _P, i]
        _cpu.cpu_depth.restype = i
        _cpu.cpu_max_threads.restype = i
    return _cpu


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def census(image):
    image = _c(image, np.uint8)
    h, w = image.shape
    out = np.empty((h, w), np.uint64)
    lib().orc_census(_ptr(image), w, h, _ptr(out))
    return out


def costs(fl, fr, disparities):
    fl, fr = _c(fl, np.uint64), _c(fr, np.uint64)
    h, w = fl.shape
    out = np.empty((h, w, disparities), np.uint8)
    lib().orc_costs(_ptr(fl), _ptr(fr), w, h, disparities, _ptr(out))
    return out


def path(costs_volume, dy, dx, p1, p2):
    c = _c(costs_volume, np.uint8)
    h, w, d = c.shape
    out = np.empty((h, w, d), np.int32)
    lib().orc_path(_ptr(c), w, h, d, dy, dx, p1, p2, _ptr(out))
    return out


def aggregate(costs_volume, paths, p1, p2):
    c = _c(costs_volume, np.uint8)
    h, w, d = c.shape
    out = np.empty((h, w, d), np.uint16)
    lib().orc_aggregate(_ptr(c), w, h, d, paths, p1, p2, _ptr(out))
    return out


def select(sums, uniqueness):
    s = _c(sums, np.uint16)
    h, w, d = s.shape
    disp = np.empty((h, w), np.float32)
    keys = np.empty((h, w), np.uint32)
    lib().orc_select(_ptr(s), w, h, d, float(np.float32(uniqueness)), _ptr(disp), _ptr(keys))
    return disp, keys


def consistency(disparity, keys, disparities):
    disparity, keys = _c(disparity, np.float32), _c(keys, np.uint32)
    h, w = disparity.shape
    out = np.empty((h, w), np.float32)
    lib().orc_consistency(_ptr(disparity), _ptr(keys), w, h, disparities, _ptr(out))
    return out


def median3(image):
    image = _c(image, np.float32)
    h, w = image.shape
    out = np.empty((h, w), np.float32)
    lib().orc_median3(_ptr(image), w, h, _ptr(out))
    return out


def median(image, kernel_size):
    """filters.rs:134-216: the middle of each pixel's kernel_size^2 window (3 or 5)."""
    image = _c(image, np.float32)
    h, w = image.shape
    out = np.empty((h, w), np.float32)
    lib().orc_median(_ptr(image), w, h, int(kernel_size), _ptr(out))
    return out


def sharpness_sums(left, right):
    """measures.rs:11-51: [left energy, right energy, left grey, right grey, left pairs,
    right pairs]."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    h, w = left.shape
    out = np.empty(6, np.float32)
    lib().orc_sharpness(_ptr(left), _ptr(right), w, h, _ptr(out))
    return out


def alignment_sums(left, right, disparity, bands=4, reach=2, row_step=4):
    """measures.rs:53-108: bands rows of 2 reach + 2 sums."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    disparity = _c(disparity, np.float32)
    h, w = left.shape
    out = np.empty((bands, 2 * reach + 2), np.float32)
    lib().orc_alignment(_ptr(left), _ptr(right), _ptr(disparity), w, h, bands, reach, row_step,
                        _ptr(out))
    return out


def sgm(left, right, disparities=64, paths=8, p1=8, p2=64, uniqueness=1.0526315789473684):
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    h, w = left.shape
    out = np.empty((h, w), np.float32)
    rc = lib().orc_sgm(_ptr(left), _ptr(right), w, h, disparities, paths, p1, p2,
                       float(np.float32(uniqueness)), _ptr(out))
    if rc:
        raise MemoryError("oracle sgm ran out of memory")
    return out


def sgm_raw(left, right, disparities=64, paths=8, p1=8, p2=64, uniqueness=1.0526315789473684):
    """sgm/3's own output (nx_stereo.ex:475-480): a match at 0 stays 0."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    h, w = left.shape
    out = np.empty((h, w), np.float32)
    rc = lib().orc_sgm_raw(_ptr(left), _ptr(right), w, h, disparities, paths, p1, p2,
                           float(np.float32(uniqueness)), _ptr(out))
    if rc:
        raise MemoryError("oracle sgm ran out of memory")
    return out


def resize_area(frame, size):
    """image.ex:402-476: the area resize of a u8 frame to size = (height, width)."""
    frame = _c(frame, np.uint8)
    fh, fw = frame.shape
    h, w = size
    out = np.empty((h, w), np.uint8)
    lib().orc_resize_area(_ptr(frame), fw, fh, w, h, _ptr(out))
    return out


# Census windows of sgm_band/4 (nx_stereo.ex:28), by the index the C code takes.
CENSUS_WINDOWS = ("sparse_9x7", "dense_9x5", "dense_9x5_blend")


def shrink(image):
    """nx_stereo.ex:236-247: the rounded 2x2 average."""
    image = _c(image, np.uint8)
    h, w = image.shape
    out = np.empty((h // 2, w // 2), np.uint8)
    lib().orc_shrink(_ptr(image), w, h, _ptr(out))
    return out


def census_window(image, window="sparse_9x7"):
    image = _c(image, np.uint8)
    h, w = image.shape
    out = np.empty((h, w), np.uint64)
    lib().orc_census_window(_ptr(image), w, h, CENSUS_WINDOWS.index(window), _ptr(out))
    return out


def band_base(coarse, band=16, margin=4):
    """nx_stereo.ex:249-259: coarse f32 {h, w} -> base s32 {2h, 2w}."""
    coarse = _c(coarse, np.float32)
    ch, cw = coarse.shape
    out = np.empty((2 * ch, 2 * cw), np.int32)
    lib().orc_band_base(_ptr(coarse), cw, ch, band, margin, _ptr(out))
    return out


def band_costs(fl, fr, searched, band):
    fl, fr, searched = _c(fl, np.uint64), _c(fr, np.uint64), _c(searched, np.int32)
    h, w = fl.shape
    out = np.empty((h, w, band), np.uint8)
    lib().orc_band_costs(_ptr(fl), _ptr(fr), _ptr(searched), w, h, band, _ptr(out))
    return out


def band_path(costs_volume, searched, dy, dx, p1, p2):
    c, searched = _c(costs_volume, np.uint8), _c(searched, np.int32)
    h, w, band = c.shape
    out = np.empty((h, w, band), np.int32)
    lib().orc_band_path(_ptr(c), _ptr(searched), w, h, band, dy, dx, p1, p2, _ptr(out))
    return out


def band_aggregate(costs_volume, searched, paths, p1, p2):
    c, searched = _c(costs_volume, np.uint8), _c(searched, np.int32)
    h, w, band = c.shape
    out = np.empty((h, w, band), np.uint16)
    lib().orc_band_aggregate(_ptr(c), _ptr(searched), w, h, band, paths, p1, p2, _ptr(out))
    return out


def band_select(sums, base, disparities, uniqueness):
    s, base = _c(sums, np.uint16), _c(base, np.int32)
    h, w, band = s.shape
    disp = np.empty((h, w), np.float32)
    keys = np.empty((h, w), np.uint32)
    lib().orc_band_select(_ptr(s), _ptr(base), w, h, band, disparities,
                          float(np.float32(uniqueness)), _ptr(disp), _ptr(keys))
    return disp, keys


def sgm_band(left, right, base, disparities=64, band=16, paths=8, census="sparse_9x7", p1=8,
             p2=64, uniqueness=1.0526315789473684):
    """sgm_band/4 (nx_stereo.ex:286-322)."""
    left, right, base = _c(left, np.uint8), _c(right, np.uint8), _c(base, np.int32)
    h, w = left.shape
    out = np.empty((h, w), np.float32)
    rc = lib().orc_sgm_band(_ptr(left), _ptr(right), _ptr(base), w, h, disparities, band, paths,
                            CENSUS_WINDOWS.index(census), p1, p2,
                            float(np.float32(uniqueness)), _ptr(out))
    if rc:
        raise MemoryError("oracle sgm_band ran out of memory")
    return out


def coarse_to_fine(left, right, band=16, margin=4, paths=8, coarse_paths=None,
                   fine_census="sparse_9x7", p1=8, p2=64, uniqueness=1.0526315789473684,
                   positive=False):
    """sgm_coarse_to_fine/3 (nx_stereo.ex:164-214); positive: the Disparity stage's output
    (stage/disparity.ex:162), where a match at 0 is unmatched."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    h, w = left.shape
    out = np.empty((h, w), np.float32)
    rc = lib().orc_coarse_to_fine(_ptr(left), _ptr(right), w, h, band, margin, paths,
                                  coarse_paths or paths, CENSUS_WINDOWS.index(fine_census), p1,
                                  p2, float(np.float32(uniqueness)), 1 if positive else 0,
                                  _ptr(out))
    if rc:
        raise MemoryError("oracle coarse_to_fine ran out of memory")
    return out


def rectify(frame, k, d, inverse_pr, size):
    frame = _c(frame, np.uint8)
    fh, fw = frame.shape
    height, width = size
    k, d, ipr = _c(k, np.float32), _c(d, np.float32), _c(inverse_pr, np.float32)
    out = np.empty((height, width), np.uint8)
    lib().orc_rectify(_ptr(frame), fw, fh, _ptr(k), _ptr(d), _ptr(ipr), width, height, _ptr(out))
    return out


def despeckle(disparity, min_support, max_diff, radius=3):
    disparity = _c(disparity, np.float32)
    h, w = disparity.shape
    out = np.empty((h, w), np.float32)
    lib().orc_despeckle(_ptr(disparity), w, h, float(np.float32(min_support)),
                        float(np.float32(max_diff)), radius, _ptr(out))
    return out


def depth(disparity, focal_length_px, baseline_m):
    disparity = _c(disparity, np.float32)
    mm = np.empty(disparity.shape, np.uint16)
    m = np.empty(disparity.shape, np.float32)
    lib().orc_depth(_ptr(disparity), disparity.size, float(np.float32(focal_length_px)),
                    float(np.float32(baseline_m)), _ptr(mm), _ptr(m))
    return mm, m


def cloud(depth_m, decimation, focal_length_px, cx, cy, nearest=0.05, farthest=20.0):
    depth_m = _c(depth_m, np.float32)
    h, w = depth_m.shape
    points = np.empty((h // decimation, w // decimation, 3), np.float32)
    kept = lib().orc_cloud(_ptr(depth_m), w, h, decimation, *(float(np.float32(v)) for v in (
        focal_length_px, cx, cy, nearest, farthest)), _ptr(points))
    return points, kept


class Temporal:
    """filters.rs:270-319: the carried previous disparity and missed counts."""

    def __init__(self, shape, max_missed=255.0):
        self.previous = np.full(shape, -1.0, np.float32)
        self.missed = np.zeros(shape, np.float32)
        self.max_missed = max_missed

    def filter(self, disparity, alpha, delta, hold):
        out = _c(disparity, np.float32).copy()
        lib().orc_temporal(_ptr(out), _ptr(self.previous), _ptr(self.missed), out.size,
                           *(float(np.float32(v)) for v in (alpha, delta, hold, self.max_missed)))
        return out


class Quality:
    """measures.rs:110-155: the carried previous disparity."""

    def __init__(self, shape):
        self.previous = np.full(shape, -1.0, np.float32)

    def sums(self, disparity):
        disparity = _c(disparity, np.float32)
        h, w = disparity.shape
        out = np.empty(5, np.float32)
        lib().orc_quality(_ptr(disparity), _ptr(self.previous), w, h, _ptr(out))
        return out


def cpu_sgm(left, right, disparities=64, paths=8, p1=8, p2=64, uniqueness=1.0526315789473684,
            threads=0):
    """The multi-threaded stereo_match-structured baseline; threads=0 takes every core."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    h, w = left.shape
    out = np.empty((h, w), np.float32)
    rc = cpu().cpu_sgm(_ptr(left), _ptr(right), w, h, disparities, paths, p1, p2,
                       float(np.float32(uniqueness)), _ptr(out), threads)
    if rc:
        raise MemoryError("cpu baseline ran out of memory")
    return out


def cpu_rectify_pair(left_frame, right_frame, left_camera, right_camera, size, threads=0):
    """The multi-threaded baseline's Rectification: both frames through their cameras' maps
    (each built once and kept while its camera is the same), rows from a queue."""
    left_frame, right_frame = _c(left_frame, np.uint8), _c(right_frame, np.uint8)
    fh, fw = left_frame.shape
    h, w = size
    outs = np.empty((h, w), np.uint8), np.empty((h, w), np.uint8)
    cams = [[_c(cam[key], np.float32) for key in ("k", "d", "inverse_pr")]
            for cam in (left_camera, right_camera)]
    rc = cpu().cpu_rectify_pair(_ptr(left_frame), _ptr(right_frame), fw, fh,
                                *[_ptr(a) for a in cams[0] + cams[1]], w, h, _ptr(outs[0]),
                                _ptr(outs[1]), threads)
    if rc:
        raise MemoryError("cpu baseline ran out of memory")
    return outs


def cpu_despeckle(disparity, min_support, max_diff, radius=3, threads=0):
    """The multi-threaded baseline's SpeckleFilter, rows from a queue."""
    disparity = _c(disparity, np.float32)
    h, w = disparity.shape
    out = np.empty_like(disparity)
    rc = cpu().cpu_despeckle(_ptr(disparity), w, h, float(min_support), float(max_diff), radius,
                             _ptr(out), threads)
    if rc:
        raise MemoryError("cpu baseline ran out of memory")
    return out


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), for the baseline's report."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_depth(disparity, focal_length_px, baseline_m, threads=0):
    """The multi-threaded baseline's Depth, rows from a queue: (depth_mm, depth_m)."""
    disparity = _c(disparity, np.float32)
    h, w = disparity.shape
    mm, m = np.empty((h, w), np.uint16), np.empty((h, w), np.float32)
    cpu().cpu_depth(_ptr(disparity), w, h, float(np.float32(focal_length_px)),
                    float(np.float32(baseline_m)), _ptr(mm), _ptr(m), threads)
    return mm, m
