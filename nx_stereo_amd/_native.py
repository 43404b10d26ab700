"""ctypes bindings of libnxs_hip.so (include/nxs_hip.h).

The library must be built (``python -m nx_stereo_amd.build``); there is no
fallback. Loading fails loudly when it is missing, and every entry point that
reaches the GPU fails loudly when the HIP runtime finds no device.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libnxs_hip.so")

NXS_OK, NXS_TIMEOUT, NXS_EINVAL, NXS_EDEVICE, NXS_ENOMEM, NXS_EUNKNOWN = 0, 1, -1, -2, -3, -4

# Every symbol include/nxs_hip.h declares.
SYMBOLS = (
    "nxs_open", "nxs_close", "nxs_submit", "nxs_wait", "nxs_reply", "nxs_release_reply",
    "nxs_poll", "nxs_set_hbm_budget", "nxs_stats", "nxs_coalescing", "nxs_match_workspace_size",
    "nxs_match_device",
    "nxs_match_device_noalloc", "nxs_match_host",
    "nxs_band_workspace_size", "nxs_match_band_device", "nxs_match_band_host",
    "nxs_coarse_to_fine_workspace_size", "nxs_coarse_to_fine_device", "nxs_coarse_to_fine_host",
    "nxs_run_device", "nxs_profile_start", "nxs_profile_stop", "nxs_schedule_set",
    "nxs_schedule_get", "nxs_batch_create", "nxs_batch_create_planned",
    "nxs_batch_run", "nxs_batch_describe", "nxs_batch_destroy", "nxs_last_error",
)

DONE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64,
                           ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)

_lib = None


class NxsTensor(ctypes.Structure):
    """struct nxs_tensor: a named device tensor."""

    _fields_ = [("name", ctypes.c_char_p), ("type", ctypes.c_uint8), ("rank", ctypes.c_uint8),
                ("shape", ctypes.c_uint32 * 4), ("data", ctypes.c_void_p)]


class NxsPair(ctypes.Structure):
    """struct nxs_pair: one pair of a batch, its images and output in device memory."""

    _fields_ = [("left", ctypes.c_void_p), ("right", ctypes.c_void_p),
                ("disparity", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("disparities", ctypes.c_int), ("paths", ctypes.c_int), ("p1", ctypes.c_int),
                ("p2", ctypes.c_int), ("uniqueness", ctypes.c_float)]


class NxsSchedule(ctypes.Structure):
    """struct nxs_schedule: how matcher launches of one size aggregate."""

    _fields_ = [("prefetch", ctypes.c_int), ("split", ctypes.c_int), ("lanes", ctypes.c_int),
                ("waves", ctypes.c_int), ("meet", ctypes.c_int)]


class NxsBatchPlan(ctypes.Structure):
    """struct nxs_batch_plan: a batch's branches, pairs a launch, fused checkpoints and
    whether the fused aggregation runs in persistent blocks on a work queue."""

    _fields_ = [("branches", ctypes.c_int), ("chunk", ctypes.c_int), ("fused", ctypes.c_int),
                ("queue", ctypes.c_int)]


class NativeError(RuntimeError):
    """A failing call into libnxs_hip, with the library's reason."""


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise NativeError(f"{LIBRARY} is not built: run `python -m nx_stereo_amd.build`")
    lib = ctypes.CDLL(LIBRARY)
    P, i, f, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    lib.nxs_open.argtypes = [i, ctypes.POINTER(P), ctypes.c_char_p, sz]
    lib.nxs_open.restype = i
    lib.nxs_close.argtypes = [P]
    lib.nxs_close.restype = None
    lib.nxs_submit.argtypes = [P, P, sz, DONE_FN, P]
    lib.nxs_submit.restype = ctypes.c_int64
    lib.nxs_wait.argtypes = [P, ctypes.c_int64, i]
    lib.nxs_wait.restype = i
    lib.nxs_reply.argtypes = [P, ctypes.c_int64, ctypes.POINTER(P), ctypes.POINTER(sz)]
    lib.nxs_reply.restype = i
    lib.nxs_release_reply.argtypes = [P, ctypes.c_int64]
    lib.nxs_release_reply.restype = None
    lib.nxs_poll.argtypes = [P, i]
    lib.nxs_poll.restype = i
    lib.nxs_set_hbm_budget.argtypes = [P, sz]
    lib.nxs_set_hbm_budget.restype = i
    lib.nxs_stats.argtypes = [P, ctypes.POINTER(sz), ctypes.POINTER(i)]
    lib.nxs_stats.restype = i
    lib.nxs_coalescing.argtypes = [P, i, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    lib.nxs_coalescing.restype = i
    lib.nxs_match_workspace_size.argtypes = [i, i, i, i]
    lib.nxs_match_workspace_size.restype = sz
    lib.nxs_match_device.argtypes = [P, P, i, i, i, i, i, i, f, P, P, sz, P]
    lib.nxs_match_device.restype = i
    lib.nxs_match_device_noalloc.argtypes = [P, P, i, i, i, i, i, i, f, P, P]
    lib.nxs_match_device_noalloc.restype = i
    lib.nxs_match_host.argtypes = [P, P, i, i, i, i, i, i, f] + [P] * 8
    lib.nxs_match_host.restype = i
    lib.nxs_band_workspace_size.argtypes = [i, i, i, i]
    lib.nxs_band_workspace_size.restype = sz
    lib.nxs_match_band_device.argtypes = [P, P, P, i, i, i, i, i, i, i, i, f, P, P, sz, P]
    lib.nxs_match_band_device.restype = i
    lib.nxs_match_band_host.argtypes = [P, P, P, i, i, i, i, i, i, i, i, f] + [P] * 8
    lib.nxs_match_band_host.restype = i
    lib.nxs_coarse_to_fine_workspace_size.argtypes = [i, i, i, i, i]
    lib.nxs_coarse_to_fine_workspace_size.restype = sz
    lib.nxs_coarse_to_fine_device.argtypes = [P, P, i, i, i, i, i, i, i, i, i, f, P, P, sz, P]
    lib.nxs_coarse_to_fine_device.restype = i
    lib.nxs_coarse_to_fine_host.argtypes = [P, P, i, i, i, i, i, i, i, i, i, f] + [P] * 5
    lib.nxs_coarse_to_fine_host.restype = i
    lib.nxs_run_device.argtypes = [P, P, sz, ctypes.POINTER(NxsTensor), i,
                                   ctypes.POINTER(NxsTensor), i, ctypes.POINTER(i), P]
    lib.nxs_run_device.restype = i
    lib.nxs_profile_start.argtypes = [i]
    lib.nxs_profile_start.restype = i
    lib.nxs_profile_stop.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i), i]
    lib.nxs_profile_stop.restype = i
    lib.nxs_schedule_set.argtypes = [i, i, i, i, ctypes.POINTER(NxsSchedule)]
    lib.nxs_schedule_set.restype = i
    lib.nxs_schedule_get.argtypes = [i, i, i, i, ctypes.POINTER(NxsSchedule)]
    lib.nxs_schedule_get.restype = i
    lib.nxs_batch_create.argtypes = [ctypes.POINTER(NxsPair), i, i, ctypes.POINTER(P)]
    lib.nxs_batch_create.restype = i
    lib.nxs_batch_create_planned.argtypes = [ctypes.POINTER(NxsPair), i,
                                             ctypes.POINTER(NxsBatchPlan), ctypes.POINTER(P)]
    lib.nxs_batch_create_planned.restype = i
    lib.nxs_batch_run.argtypes = [P, P]
    lib.nxs_batch_run.restype = i
    lib.nxs_batch_describe.argtypes = [P]
    lib.nxs_batch_describe.restype = ctypes.c_char_p
    lib.nxs_batch_destroy.argtypes = [P]
    lib.nxs_batch_destroy.restype = None
    lib.nxs_last_error.argtypes = []
    lib.nxs_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def last_error() -> str:
    return load().nxs_last_error().decode("utf-8", "replace")


def _check(rc: int, what: str) -> None:
    if rc != NXS_OK:
        raise NativeError(f"{what} failed ({rc}): {last_error()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def match_host(left: np.ndarray, right: np.ndarray, disparities: int = 64, paths: int = 8,
               p1: int = 8, p2: int = 64, uniqueness: float = 1.0526315789473684,
               intermediates: bool = False) -> dict:
    """Stage.Disparity on the GPU from host images; with `intermediates`, every product."""
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    h, w = left.shape
    out = {"disparity": np.empty((h, w), np.float32)}
    if intermediates:
        out.update(
            census_left=np.empty((h, w), np.uint64), census_right=np.empty((h, w), np.uint64),
            costs=np.empty((h, w, disparities), np.uint8),
            sums=np.empty((h, w, disparities), np.uint16),
            selected=np.empty((h, w), np.float32), keys=np.empty((h, w), np.uint32),
            checked=np.empty((h, w), np.float32))
    g = out.get
    rc = load().nxs_match_host(
        _ptr(left), _ptr(right), w, h, disparities, paths, p1, p2,
        float(np.float32(uniqueness)), _ptr(g("census_left")), _ptr(g("census_right")),
        _ptr(g("costs")), _ptr(g("sums")), _ptr(g("selected")), _ptr(g("keys")),
        _ptr(g("checked")), _ptr(out["disparity"]))
    _check(rc, "nxs_match_host")
    return out


def workspace_size(width: int, height: int, disparities: int, paths: int) -> int:
    return int(load().nxs_match_workspace_size(width, height, disparities, paths))


def match_device(left_ptr: int, right_ptr: int, width: int, height: int, disparities: int,
                 paths: int, p1: int, p2: int, uniqueness: float, out_ptr: int,
                 workspace_ptr: int, workspace_bytes: int, stream: int = 0) -> None:
    """Stage.Disparity on device buffers (raw device addresses), enqueued on `stream`."""
    rc = load().nxs_match_device(
        ctypes.c_void_p(left_ptr), ctypes.c_void_p(right_ptr), width, height, disparities,
        paths, p1, p2, float(np.float32(uniqueness)), ctypes.c_void_p(out_ptr),
        ctypes.c_void_p(workspace_ptr), workspace_bytes, ctypes.c_void_p(stream or None))
    _check(rc, "nxs_match_device")


# nxs_match_device_noalloc's scratch (NXS_NOALLOC_SCRATCH_BYTES in include/nxs_hip.h).
NOALLOC_SCRATCH_BYTES = 64 << 20


def match_device_noalloc(left_ptr: int, right_ptr: int, width: int, height: int,
                         disparities: int, paths: int, p1: int, p2: int, uniqueness: float,
                         out_ptr: int, stream: int = 0) -> None:
    """Stage.Disparity on device buffers with the library's scratch as workspace (nothing
    allocated), enqueued on `stream`."""
    rc = load().nxs_match_device_noalloc(
        ctypes.c_void_p(left_ptr), ctypes.c_void_p(right_ptr), width, height, disparities,
        paths, p1, p2, float(np.float32(uniqueness)), ctypes.c_void_p(out_ptr),
        ctypes.c_void_p(stream or None))
    _check(rc, "nxs_match_device_noalloc")


# sgm_band/4's census windows (nx_stereo.ex:28), by the index the C ABI takes.
CENSUS_WINDOWS = ("sparse_9x7", "dense_9x5", "dense_9x5_blend")


def _census_index(window: str, key: str) -> int:
    if window not in CENSUS_WINDOWS:
        raise ValueError(f":{key} must be one of {list(CENSUS_WINDOWS)}, got {window!r}")
    return CENSUS_WINDOWS.index(window)


def sgm_band(left: np.ndarray, right: np.ndarray, base: np.ndarray, disparities: int = 64,
             band: int = 16, paths: int = 8, census: str = "sparse_9x7", p1: int = 8,
             p2: int = 64, uniqueness: float = 1.0526315789473684,
             intermediates: bool = False) -> dict:
    """NxStereo.sgm_band/4 (nx_stereo.ex:262-322) on the GPU from host arrays: each pixel
    searches [base, base + band); a negative base gives no match. With `intermediates`,
    every product of nxs_match_band_host."""
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    h, w = left.shape
    if band not in (8, 16, 32, 64):
        raise ValueError(f":band must be 8, 16, 32 or 64, got {band!r}")
    base = np.asarray(base)
    if base.dtype != np.int32 or base.shape != (h, w):
        raise ValueError(f"sgm_band expects base as an s32 tensor of shape {(h, w)}, "
                         f"got {base.dtype} {base.shape}")
    base = np.ascontiguousarray(base)
    index = _census_index(census, "census")
    out = {"disparity": np.empty((h, w), np.float32)}
    if intermediates:
        out.update(
            census_left=np.empty((h, w), np.uint64), census_right=np.empty((h, w), np.uint64),
            costs=np.empty((h, w, band), np.uint8), sums=np.empty((h, w, band), np.uint16),
            selected=np.empty((h, w), np.float32), keys=np.empty((h, w), np.uint32),
            checked=np.empty((h, w), np.float32))
    g = out.get
    rc = load().nxs_match_band_host(
        _ptr(left), _ptr(right), _ptr(base), w, h, disparities, band, paths, index, p1, p2,
        float(np.float32(uniqueness)), _ptr(g("census_left")), _ptr(g("census_right")),
        _ptr(g("costs")), _ptr(g("sums")), _ptr(g("selected")), _ptr(g("keys")),
        _ptr(g("checked")), _ptr(out["disparity"]))
    _check(rc, "nxs_match_band_host")
    return out


def sgm_coarse_to_fine(left: np.ndarray, right: np.ndarray, disparities: int = 128,
                       band: int = 16, margin: int = 4, paths: int = 8,
                       coarse_paths: int | None = None, fine_census: str = "sparse_9x7",
                       p1: int = 8, p2: int = 64, uniqueness: float = 1.0526315789473684,
                       intermediates: bool = False) -> dict:
    """NxStereo.sgm_coarse_to_fine/3 (nx_stereo.ex:164-259) on the GPU from host arrays.
    With `intermediates`, the half images, the coarse map and the bases as well."""
    left = np.ascontiguousarray(left, np.uint8)
    right = np.ascontiguousarray(right, np.uint8)
    h, w = left.shape
    if h % 2 or w % 2:
        raise ValueError("sgm_coarse_to_fine expects images of even height and width, "
                         f"got {(h, w)}")
    if disparities != 128:
        raise ValueError(":disparities must be 128: the coarse matching takes 64, "
                         f"got {disparities!r}")
    index = _census_index(fine_census, "fine_census")
    out = {"disparity": np.empty((h, w), np.float32)}
    if intermediates:
        out.update(half_left=np.empty((h // 2, w // 2), np.uint8),
                   half_right=np.empty((h // 2, w // 2), np.uint8),
                   coarse=np.empty((h // 2, w // 2), np.float32),
                   base=np.empty((h, w), np.int32))
    g = out.get
    rc = load().nxs_coarse_to_fine_host(
        _ptr(left), _ptr(right), w, h, band, margin, paths, coarse_paths or paths, index, p1, p2,
        float(np.float32(uniqueness)), _ptr(g("half_left")), _ptr(g("half_right")),
        _ptr(g("coarse")), _ptr(g("base")), _ptr(out["disparity"]))
    _check(rc, "nxs_coarse_to_fine_host")
    return out


def band_workspace_size(width: int, height: int, band: int, paths: int) -> int:
    return int(load().nxs_band_workspace_size(width, height, band, paths))


def coarse_to_fine_workspace_size(width: int, height: int, band: int, paths: int,
                                  coarse_paths: int) -> int:
    return int(load().nxs_coarse_to_fine_workspace_size(width, height, band, paths, coarse_paths))


def coarse_to_fine_device(left_ptr: int, right_ptr: int, width: int, height: int, band: int,
                          margin: int, paths: int, coarse_paths: int, fine_census: int, p1: int,
                          p2: int, uniqueness: float, out_ptr: int, workspace_ptr: int,
                          workspace_bytes: int, stream: int = 0) -> None:
    """sgm_coarse_to_fine/3 on device buffers (raw device addresses), enqueued on `stream`."""
    rc = load().nxs_coarse_to_fine_device(
        ctypes.c_void_p(left_ptr), ctypes.c_void_p(right_ptr), width, height, band, margin,
        paths, coarse_paths, fine_census, p1, p2, float(np.float32(uniqueness)),
        ctypes.c_void_p(out_ptr), ctypes.c_void_p(workspace_ptr), workspace_bytes,
        ctypes.c_void_p(stream or None))
    _check(rc, "nxs_coarse_to_fine_device")


def set_schedule(width, height, disparities, paths, prefetch, split, lanes=16, waves=1,
                 meet=0) -> None:
    s = NxsSchedule(prefetch, int(split), lanes, waves, meet)
    _check(load().nxs_schedule_set(width, height, disparities, paths, ctypes.byref(s)),
           "nxs_schedule_set")


def get_schedule(width, height, disparities, paths) -> tuple[int, bool, int, int, int]:
    s = NxsSchedule()
    _check(load().nxs_schedule_get(width, height, disparities, paths, ctypes.byref(s)),
           "nxs_schedule_get")
    return s.prefetch, bool(s.split), s.lanes, s.waves, s.meet


BATCH_DEFAULT, BATCH_TUNED, BATCH_FUSED = 0, 1, 2  # nxs_batch_create's modes


class Batch:
    """nxs_batch: pairs (left, right, out device addresses; width, height, disparities,
    paths, p1, p2, uniqueness) captured into one graph; run() replays it on a stream.
    autotune: the class schedules and plans timed on the batch (NXS_BATCH_TUNED); fused:
    every class in one launch per kernel (NXS_BATCH_FUSED); plan: (branches, chunk, fused
    [, queue]) pinned, untuned (nxs_batch_create_planned)."""

    def __init__(self, pairs, autotune: bool = True, fused: bool = False, plan=None):
        lib = load()
        arr = (NxsPair * len(pairs))()
        for k, (left, right, out, w, h, d, paths, p1, p2, u) in enumerate(pairs):
            arr[k] = NxsPair(left, right, out, w, h, d, paths, p1, p2, float(np.float32(u)))
        handle = ctypes.c_void_p()
        if plan is not None:
            pinned = NxsBatchPlan(*plan)
            _check(lib.nxs_batch_create_planned(arr, len(pairs), ctypes.byref(pinned),
                                                ctypes.byref(handle)), "nxs_batch_create_planned")
        else:
            mode = BATCH_FUSED if fused else (BATCH_TUNED if autotune else BATCH_DEFAULT)
            _check(lib.nxs_batch_create(arr, len(pairs), mode, ctypes.byref(handle)),
                   "nxs_batch_create")
        self.lib, self.handle = lib, handle

    def run(self, stream: int = 0) -> None:
        _check(self.lib.nxs_batch_run(self.handle, ctypes.c_void_p(stream or None)), "nxs_batch_run")

    def describe(self) -> str:
        return self.lib.nxs_batch_describe(self.handle).decode()

    def close(self) -> None:
        if self.handle:
            self.lib.nxs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
