"""The stages HipSegment computes, with the settings API of NxStereo.Stage.

Each class mirrors its Elixir stage module (lib/nx_stereo/stage/*.ex): the
declared settings (default, type, range, values, greater_than, live), init
with NxStereo.Stage.init_settings/2's checks and messages
(lib/nx_stereo/stage.ex:153-212), params (the live inputs of one pair, sent
as f32), native_kind, native_settings and native_constants (what the
configure record carries, native_segment.ex:256-290).

Geometry is a dict: width, height, frame_width, frame_height,
focal_length_px, baseline_m, principal_point (cx, cy), and left_camera /
right_camera, each {"k": 9, "d": 5, "inverse_pr": 9} f32 values.
"""
from __future__ import annotations

import math

import numpy as np


class StageError(ValueError):
    """A setting the stage refuses, with NxStereo.Stage's message."""


def _inspect(value):
    if isinstance(value, str):
        return f'"{value}"'
    if value is None:
        return "nil"
    if isinstance(value, bool):
        return "true" if value else "false"
    return repr(value)


def _typed(kind, value):
    if kind == "integer":
        return isinstance(value, int) and not isinstance(value, bool)
    if kind == "float":
        return isinstance(value, (int, float)) and not isinstance(value, bool)
    if kind == "string":
        return isinstance(value, str)
    if kind == "boolean":
        return isinstance(value, bool)
    return True


def _accepts(setting, value):
    if "values" in setting:
        return value in setting["values"]
    if "range" in setting:
        lo, hi = setting["range"]
        if not (_typed(setting["type"], value) and lo <= value <= hi):
            return False
        if "step" in setting:
            return round((value - lo) * 1_000_000) % round(setting["step"] * 1_000_000) == 0
        return True
    return _typed(setting["type"], value)


def _accepted(setting):
    if "values" in setting:
        return "one of [" + ", ".join(_inspect(v) for v in setting["values"]) + "]"
    if "range" in setting:
        lo, hi = setting["range"]
        return f"from {lo} to {hi}"
    return f"of type {setting['type']}"


class Stage:
    """NxStereo.Stage's defaults: no settings, no params, no carry."""

    SETTINGS: dict = {}
    KIND: str = ""
    OUTPUTS: tuple = ()

    def __init__(self, **settings):
        self.state = self.init(settings)

    @classmethod
    def init_settings(cls, settings: dict) -> dict:
        unknown = [k for k in settings if k not in cls.SETTINGS]
        if unknown:
            raise StageError(f"unknown settings {unknown!r}")
        values = {k: s["default"] for k, s in cls.SETTINGS.items()}
        values.update(settings)
        for key, setting in cls.SETTINGS.items():
            value = values[key]
            if value is None and setting["default"] is None:
                continue
            if not _accepts(setting, value):
                raise StageError(f"{key} must be {_accepted(setting)}, got {_inspect(value)}")
        for key, setting in cls.SETTINGS.items():
            other = setting.get("greater_than")
            if other and values[key] <= values[other]:
                raise StageError(f"{key} must be above {other} ({values[other]}), got {values[key]}")
        return values

    def init(self, settings):
        return self.init_settings(settings)

    def params(self, geometry) -> dict:
        return {}

    def native_settings(self) -> list:
        return []

    def native_constants(self) -> dict:
        return {}

    def compute(self):
        return None


class Decode(Stage):
    """stage/decode.ex: both frames area-resized (image.ex:402-476) to the working size, for
    frames that are already rectified; u8 frames, or frames packed as u32 words."""

    KIND = "decode"
    OUTPUTS = ("left", "right")


class Rectification(Stage):
    """stage/rectification.ex: both frames undistorted, rectified and shrunk."""

    KIND = "rectification"
    OUTPUTS = ("left", "right")
    SETTINGS = {"right_row_offset": {"default": 0.0, "type": "float", "range": (-8, 8), "live": True}}

    def params(self, geometry):
        right = dict(geometry["right_camera"])
        right["inverse_pr"] = move_rows(right["inverse_pr"], self.state["right_row_offset"])
        return {"left": geometry["left_camera"], "right": right}


def move_rows(inverse_pr, rows):
    """stage/rectification.ex:84-89: column 2 += rows * column 1, in f32."""
    inverse = np.asarray(inverse_pr, np.float32).reshape(3, 3).copy()
    inverse[:, 2] = inverse[:, 2] + np.float32(rows) * inverse[:, 1]
    return inverse.reshape(-1)


class Disparity(Stage):
    """stage/disparity.ex: SGM, full search or coarse-to-fine (sgm_coarse_to_fine/3). 192
    and 256 disparities extend the reference's full search. HipSegment computes
    coarse_to_fine itself; the reference's native matcher refuses it
    (stage/disparity.ex:166, 184-185), so its settings travel only to this backend:
    search (0 full, 1 coarse_to_fine), coarse_paths, band, margin and fine_census (0
    sparse_9x7, 1 dense_9x5, 2 dense_9x5_blend)."""

    KIND = "disparity"
    OUTPUTS = ("disparity",)
    CENSUS_WINDOWS = ("sparse_9x7", "dense_9x5", "dense_9x5_blend")
    SETTINGS = {
        "p1": {"default": 8, "type": "integer", "range": (0, 78), "live": True},
        "p2": {"default": 64, "type": "integer", "range": (1, 79), "greater_than": "p1", "live": True},
        "uniqueness": {"default": 1.0526315789473684, "type": "float", "range": (1, 3), "live": True},
        "num_disparities": {"default": 64, "type": "integer", "values": [64, 128, 192, 256]},
        "paths": {"default": 8, "type": "integer", "values": [3, 8]},
        "search": {"default": "full", "type": "string", "values": ["full", "coarse_to_fine"]},
        "coarse_paths": {"default": None, "type": "integer", "values": [3, 8]},
        "band": {"default": 16, "type": "integer", "values": [8, 16, 32, 64]},
        "margin": {"default": 4, "type": "integer", "range": (0, 63)},
        "fine_census": {"default": "sparse_9x7", "type": "atom", "values": list(CENSUS_WINDOWS)},
        "matcher": {"default": None, "type": "atom", "values": [None, "nx", "native"]},
    }

    def init(self, settings):
        # stage/disparity.ex:121-195 (the geometry's even sizes are checked at configure)
        values = self.init_settings(settings)
        values["coarse_paths"] = values["coarse_paths"] or values["paths"]
        if values["search"] == "coarse_to_fine":
            if values["num_disparities"] != 128:
                raise StageError("coarse_to_fine searches 128 disparities, got "
                                 f"{values['num_disparities']}")
            if values["margin"] >= values["band"]:
                raise StageError(f"margin must be below band ({values['band']}), "
                                 f"got {values['margin']}")
        elif values["fine_census"] != "sparse_9x7":
            raise StageError("fine_census requires coarse_to_fine search")
        return values

    def params(self, geometry):
        return {k: self.state[k] for k in ("p1", "p2", "uniqueness")}

    def native_settings(self):
        settings = [("num_disparities", self.state["num_disparities"]),
                    ("paths", self.state["paths"])]
        if self.state["search"] == "coarse_to_fine":
            settings += [("search", 1), ("coarse_paths", self.state["coarse_paths"]),
                         ("band", self.state["band"]), ("margin", self.state["margin"]),
                         ("fine_census", self.CENSUS_WINDOWS.index(self.state["fine_census"]))]
        return settings

    def compute(self):
        return self.state["matcher"]


class Sharpness(Stage):
    """stage/sharpness.ex: each rectified image's sharpness sums, from left and right."""

    KIND = "sharpness"
    OUTPUTS = ("sharpness_sums",)


class MedianFilter(Stage):
    """stage/median_filter.ex: the median of each pixel's kernel_size^2 window."""

    KIND = "median_filter"
    OUTPUTS = ("disparity",)
    SETTINGS = {"kernel_size": {"default": 5, "type": "integer", "values": [3, 5]}}

    def native_settings(self):
        return [("kernel_size", self.state["kernel_size"])]


class RowAlignment(Stage):
    """stage/row_alignment.ex: the row-offset sums, with the diagnostics' default bands,
    reach and row step as constants."""

    KIND = "row_alignment"
    OUTPUTS = ("alignment_sums",)
    BANDS, REACH, ROW_STEP = 4, 2, 4

    def native_constants(self):
        return {"bands": self.BANDS, "reach": self.REACH, "row_step": self.ROW_STEP}


class SpeckleFilter(Stage):
    """stage/speckle_filter.ex: 7x7 support count."""

    KIND = "speckle_filter"
    OUTPUTS = ("disparity",)
    RADIUS = 3
    SETTINGS = {
        "min_support": {"default": 12, "type": "integer", "range": (0, 48), "live": True},
        "max_diff": {"default": 1.0, "type": "float", "range": (0, 8), "live": True},
    }

    def params(self, geometry):
        return {k: self.state[k] for k in ("min_support", "max_diff")}

    def native_constants(self):
        return {"radius": self.RADIUS}


class TemporalFilter(Stage):
    """stage/temporal_filter.ex: carried previous disparity and missed counts."""

    KIND = "temporal_filter"
    OUTPUTS = ("disparity",)
    MAX_MISSED = 255
    SETTINGS = {
        "alpha": {"default": 0.4, "type": "float", "range": (0.05, 1), "live": True},
        "delta": {"default": 2.0, "type": "float", "range": (0, 8), "live": True},
        "hold": {"default": 2, "type": "integer", "range": (0, 10), "live": True},
    }

    def params(self, geometry):
        return {k: self.state[k] for k in ("alpha", "delta", "hold")}

    def native_constants(self):
        return {"max_missed": float(self.MAX_MISSED)}


class Depth(Stage):
    """stage/depth.ex: depth_mm (returned) and depth_m (left on the device)."""

    KIND = "depth"
    OUTPUTS = ("depth_mm",)

    def params(self, geometry):
        return {"focal_length_px": geometry["focal_length_px"], "baseline_m": geometry["baseline_m"]}


class PointCloud(Stage):
    """stage/point_cloud.ex: every decimation-th pixel within 5 cm to 20 m."""

    KIND = "point_cloud"
    OUTPUTS = ("cloud_grid", "cloud_points")
    SETTINGS = {"decimation": {"default": 4, "type": "integer", "range": (1, 16), "step": 1}}

    def init(self, settings):
        return self.init_settings(settings)

    def params(self, geometry):
        cx, cy = geometry["principal_point"]
        return {"focal_length_px": geometry["focal_length_px"], "cx": cx, "cy": cy}

    def native_settings(self):
        return [("decimation", self.state["decimation"])]

    def native_constants(self):
        return {"min_depth_m": 0.05, "max_depth_m": 20.0}


class Quality(Stage):
    """stage/quality.ex: the five disparity sums, carrying the previous disparity."""

    KIND = "quality"
    OUTPUTS = ("quality_sums",)


KINDS = {cls.KIND: cls for cls in (Decode, Rectification, Sharpness, Disparity, MedianFilter,
                                   SpeckleFilter, RowAlignment, TemporalFilter, Depth, PointCloud,
                                   Quality)}


def check_geometry(stage: Stage, geometry: dict) -> None:
    """The size checks the stage modules' init/2 make."""
    w, h = geometry["width"], geometry["height"]
    if isinstance(stage, Rectification) and (geometry["frame_width"] < w or geometry["frame_height"] < h):
        raise StageError(
            f"rectification only shrinks: {geometry['frame_width']}x{geometry['frame_height']} "
            f"frames are smaller than the {w}x{h} working size")
    if isinstance(stage, Decode) and (geometry["frame_width"] < w or geometry["frame_height"] < h):
        raise StageError(
            f"resize_area only shrinks, got {geometry['frame_height']}x{geometry['frame_width']} "
            f"to {h}x{w}")
    if isinstance(stage, PointCloud):
        d = stage.state["decimation"]
        if d > h or d > w:
            raise StageError(f"decimation must fit working size {w}x{h}, got {d}")


def flatten(params: dict, prefix: str = "") -> list[tuple[str, np.ndarray]]:
    """native_segment.ex:309-324: nested params as dotted names, f32 values, sorted."""
    out = []
    for key, value in params.items():
        name = prefix + str(key)
        if isinstance(value, dict):
            out.extend(flatten(value, name + "."))
        else:
            out.append((name, np.asarray(value, np.float32).reshape(-1)))
    return sorted(out, key=lambda kv: kv[0])


def setting_float(value) -> float:
    if value is True:
        return 1.0
    if value is False:
        return 0.0
    return float(value)


def build_settings(stage: Stage) -> list[tuple[str, float]]:
    """native_segment.ex:278-290: build settings, then the constants sorted by name."""
    build = [(k, setting_float(v)) for k, v in stage.native_settings()]
    constants = sorted((k, setting_float(v)) for k, v in stage.native_constants().items())
    return build + constants


def finite(value) -> bool:
    return isinstance(value, (int, float)) and math.isfinite(value)
