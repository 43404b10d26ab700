"""A segment geometry of the synthetic calibrated rig (nx_stereo_amd.scenes.rig) at a working
size: the calibration of frames of frame_width x frame_height, rectified to width x height."""
from nx_stereo_amd import scenes


def rig_geometry(width, height, frame_width=None, frame_height=None):
    frame_width, frame_height = frame_width or width, frame_height or height
    cal = scenes.rig(frame_width, frame_height)
    scale = width / frame_width
    geometry = {"width": width, "height": height, "frame_width": frame_width,
                "frame_height": frame_height, "focal_length_px": cal["focal_length_px"] * scale,
                "baseline_m": cal["baseline_m"],
                "principal_point": (width / 2.0, height / 2.0)}
    for side in ("left", "right"):
        c = cal[side]
        k = list(c["k"])
        # The calibration at the working size: the rectified camera matrix scales with it.
        p = [v * scale if i in (0, 2, 3, 5, 6) else v for i, v in enumerate(c["p"])]
        geometry[f"{side}_camera"] = scenes.camera(k, c["d"], c["r"], p)
    return geometry, cal
