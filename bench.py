#!/usr/bin/env python3
"""Throughput of NxStereo's full plain pipeline on one MI355X.

The workload: unrectified u8 frames 728x544 from a calibrated synthetic rig
(f = 873 px, principal point at the centre, k1 = -0.05, k2 = 0.01,
rectifying rotations 0.5 deg apart, baseline 0.2 m) looking at seeded
textured planes, through Rectification -> Disparity (128 disparities,
8 paths) -> SpeckleFilter -> Depth -> PointCloud (decimation 4) -> Quality,
as one HipSegment. A step is one pair through every stage with the frames
already in HBM and the outputs left there (nxs_run_device).

Prints one JSON line: pairs/s (and W*H*D cost evaluations/s), the
aggregation's achieved bytes/s against MI355X's 8 TB/s HBM peak (timed
live with HIP events around its kernels), and the CPU baseline (the
multi-threaded stereo_match-structured matcher plus the oracle's stages)
timed on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pairs/s + cost-evals/s at 728x544x128 (bit-exact); p99 latency under load"
WIDTH, HEIGHT, DISPARITIES, PATHS, DECIMATION = 728, 544, 128, 8, 4
HBM_PEAK_GBS = 8000.0
SEED = 20261203


def workload(seed: int = SEED):
    """The rig, its geometry and stage list, and one pair of raw frames."""
    from nx_stereo_amd import scenes, stages

    calibration = scenes.rig(WIDTH, HEIGHT)
    left_frame, right_frame, _ = scenes.raw_frames(calibration, seed)
    geometry = {
        "width": WIDTH, "height": HEIGHT, "frame_width": WIDTH, "frame_height": HEIGHT,
        "focal_length_px": calibration["focal_length_px"], "baseline_m": calibration["baseline_m"],
        "principal_point": calibration["principal_point"],
        "left_camera": scenes.camera(**calibration["left"]),
        "right_camera": scenes.camera(**calibration["right"]),
    }
    members = [
        ("rectification", stages.Rectification()),
        ("disparity", stages.Disparity(num_disparities=DISPARITIES, paths=PATHS)),
        ("speckle_filter", stages.SpeckleFilter()),
        ("depth", stages.Depth()),
        ("point_cloud", stages.PointCloud(decimation=DECIMATION)),
        ("quality", stages.Quality()),
    ]
    return geometry, members, left_frame, right_frame


def cpu_baseline(geometry, members, left_frame, right_frame, budget_s: float = 20.0) -> dict:
    """The CPU pipeline on the same pair: oracle stages around the multi-threaded matcher."""
    from oracle import oracle

    params = {name: stage.params(geometry) for name, stage in members}
    disparity = members[1][1].state
    speckle = params["speckle_filter"]
    cams = params["rectification"]
    temporal_quality = oracle.Quality((HEIGHT, WIDTH))

    def pipeline():
        left = oracle.rectify(left_frame, cams["left"]["k"], cams["left"]["d"],
                              cams["left"]["inverse_pr"], (HEIGHT, WIDTH))
        right = oracle.rectify(right_frame, cams["right"]["k"], cams["right"]["d"],
                               cams["right"]["inverse_pr"], (HEIGHT, WIDTH))
        d = oracle.cpu_sgm(left, right, DISPARITIES, PATHS, disparity["p1"], disparity["p2"],
                           disparity["uniqueness"])
        d = oracle.despeckle(d, speckle["min_support"], speckle["max_diff"])
        _, m = oracle.depth(d, geometry["focal_length_px"], geometry["baseline_m"])
        cx, cy = geometry["principal_point"]
        oracle.cloud(m, DECIMATION, geometry["focal_length_px"], cx, cy)
        temporal_quality.sums(d)

    pipeline()  # warm: buffers allocated, pages touched
    times = []
    start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - start < budget_s and len(times) < 20):
        t = time.perf_counter()
        pipeline()
        times.append(time.perf_counter() - t)
    median = float(np.median(times))
    return {
        "value": round(1.0 / median, 3),
        "unit": "pairs/s",
        "cores": int(oracle.cpu().cpu_max_threads()),
        "kind": "port",
        "sample": (f"{len(times)} pairs of the same 728x544x128 8-path pipeline on the host: the "
                   f"stereo_match-structured C matcher on every OpenMP thread, the oracle's scalar "
                   f"stages around it; median {median * 1e3:.1f} ms, min {min(times) * 1e3:.1f}, "
                   f"max {max(times) * 1e3:.1f}"),
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=200)
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one, and pairs shard across GPUs "
              f"with no exchange (--gpus {args.gpus} is not run)", file=sys.stderr)
        return 2

    from nx_stereo_amd import hip, segment

    geometry, members, left_frame, right_frame = workload()
    context = segment.Context(0)
    ok, seg = segment.HipSegment.build(members, geometry, context)
    reply = context.request(seg.configure_record())
    if reply is not True:
        raise RuntimeError(f"configure refused: {reply}")
    seg.configured_version = seg.version
    record = seg.run_record({})

    frames = [hip.Buffer.of(f) for f in (left_frame, right_frame)]
    stream = hip.Stream()
    entries = [("left_frame", 0, (HEIGHT, WIDTH), frames[0].ptr),
               ("right_frame", 0, (HEIGHT, WIDTH), frames[1].ptr)]
    hip.synchronize()

    def step():
        return context.run_device(record, entries, stream=stream.handle)

    for _ in range(args.warmup):
        step()
    hip.synchronize()

    events = [hip.Event() for _ in range(args.steps + 1)]
    segment.profile_start(capacity=8 * args.steps + 16)
    hip.synchronize()
    wall = time.perf_counter()
    events[0].record(stream)
    for k in range(args.steps):
        step()
        events[k + 1].record(stream)
    hip.synchronize()
    wall = time.perf_counter() - wall
    kernels = segment.profile_stop()
    gpu_ms = events[0].elapsed_ms(events[-1])
    per_pair = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(args.steps)])

    ms_per_step = gpu_ms / args.steps
    pairs_per_s = 1e3 / ms_per_step
    pixels = WIDTH * HEIGHT
    n = pixels * DISPARITIES

    # The aggregation: path_kernel (8 path directions) + select_kernel (winner, uniqueness,
    # subpixel), one launch each per pair. Algorithmic bytes per pair 6N + 8P.
    path_ms, path_launches = kernels["path"]
    select_ms, select_launches = kernels["select"]
    aggregation_s = (path_ms / max(path_launches, 1) + select_ms / max(select_launches, 1)) * 1e-3
    algorithmic = 6 * n + 8 * pixels
    achieved = algorithmic / aggregation_s / 1e9

    result = {
        "metric": METRIC,
        "value": round(pairs_per_s, 2),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {
            "workload": "full plain pipeline, 728x544 frames, 128 disparities, 8 paths, "
                        "decimation 4, one pair per step",
            "width": WIDTH, "height": HEIGHT, "disparities": DISPARITIES, "paths": PATHS,
            "stages": [name for name, _ in members], "form": "plain per-pair launches",
            "seed": SEED,
        },
        "cost_evals_per_s": round(pairs_per_s * n, 1),
        "latency_ms": {"p50": round(float(np.percentile(per_pair, 50)), 4),
                       "p99": round(float(np.percentile(per_pair, 99)), 4),
                       "note": "GPU time per pair, pairs enqueued back to back"},
        "wall_pairs_per_s": round(args.steps / wall, 2),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "aggregation: path_kernel + select_kernel",
            "kernel_us": {k: round(v[0] / max(v[1], 1) * 1e3, 2) for k, v in kernels.items()},
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(geometry, members, left_frame, right_frame)
    context.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
