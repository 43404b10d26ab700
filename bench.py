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
    """The same pipeline on the host, scheduled as stereo_match schedules it: the multi-
    threaded baseline's Rectification (cached maps), matcher (row queue, neighbour-progress
    column chunks), SpeckleFilter and Depth (row queues), then PointCloud and Quality on one
    thread, as stereo_match runs them. Every host thread OpenMP gives, then one thread;
    median, min and max as native/stereo_match/examples/bench.rs:43-59 reports them."""
    from oracle import oracle

    params = {name: stage.params(geometry) for name, stage in members}
    disparity = members[1][1].state
    speckle = params["speckle_filter"]
    cams = params["rectification"]
    f, b = geometry["focal_length_px"], geometry["baseline_m"]
    cx, cy = geometry["principal_point"]
    temporal_quality = oracle.Quality((HEIGHT, WIDTH))

    def pipeline(threads):
        left, right = oracle.cpu_rectify_pair(left_frame, right_frame, cams["left"],
                                              cams["right"], (HEIGHT, WIDTH), threads=threads)
        d = oracle.cpu_sgm(left, right, DISPARITIES, PATHS, disparity["p1"], disparity["p2"],
                           disparity["uniqueness"], threads=threads)
        d = oracle.cpu_despeckle(d, speckle["min_support"], speckle["max_diff"], threads=threads)
        _, m = oracle.cpu_depth(d, f, b, threads=threads)
        oracle.cloud(m, DECIMATION, f, cx, cy)
        temporal_quality.sums(d)

    def timed(threads, budget, most):
        pipeline(threads)  # warm: buffers allocated, maps built, pages touched
        times = []
        start = time.perf_counter()
        while len(times) < 3 or (time.perf_counter() - start < budget and len(times) < most):
            t = time.perf_counter()
            pipeline(threads)
            times.append(time.perf_counter() - t)
        return times

    cores = int(oracle.cpu().cpu_max_threads())
    all_threads = timed(0, budget_s * 0.6, 20)
    one_thread = timed(1, budget_s * 0.4, 20)

    def stats(times):
        return (f"median {np.median(times) * 1e3:.1f} ms, min {min(times) * 1e3:.1f}, "
                f"max {max(times) * 1e3:.1f} over {len(times)} pairs")

    median = float(np.median(all_threads))
    return {
        "value": round(1.0 / median, 3),
        "unit": "pairs/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": oracle.cpu_model(),
        "one_thread_pairs_per_s": round(1.0 / float(np.median(one_thread)), 3),
        "sample": (f"the 728x544x128 8-path pipeline on the host through the C restatement of "
                   f"stereo_match (not stereo_match itself: no Rust toolchain here), scheduled "
                   f"as it schedules: {cores} threads {stats(all_threads)}; 1 thread "
                   f"{stats(one_thread)}"),
    }


def match_pair(width, height, disparities, seed):
    """A rectified planar pair for a matcher-only leg (disparities 4 .. D-8)."""
    from nx_stereo_amd import scenes

    left, right, _ = scenes.planar_pair(width, height, seed=seed, low=4, high=disparities - 8)
    return left, right


def large_pair_leg(runs: int = 10) -> dict:
    """One 1920x1080 pair at 256 disparities, 8 paths, through the matcher alone: the HBM
    measurement (its u8 volume, 531 MB, and the aggregation's traffic exceed the caches)."""
    from nx_stereo_amd import _native, hip, segment

    w, h, d = 1920, 1080, 256
    left, right = match_pair(w, h, d, SEED + 3)
    l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
    ws = hip.Buffer(_native.workspace_size(w, h, d, 8))
    stream = hip.Stream()

    def match():
        _native.match_device(l.ptr, r.ptr, w, h, d, 8, 8, 64, 1.0526315789473684, o.ptr, ws.ptr,
                             ws.nbytes, stream.handle)

    for _ in range(2):
        match()
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(runs):
        match()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / runs
    segment.profile_start(8 * runs + 8)  # the kernels' times, in a pass of their own
    for _ in range(runs):
        match()
    stream.synchronize()
    kernels = segment.profile_stop()
    us = {k: v[0] / v[1] * 1e3 for k, v in kernels.items() if v[1]}
    pixels, n = w * h, w * h * d
    aggregation_us = us["path"] + us["select"]
    agg = (6 * n + 8 * pixels) / (aggregation_us * 1e-6) / 1e9
    # Census and costs in one kernel: the two images in, the cost volume out.
    cost = (n + 2 * pixels) / (us["census_cost"] * 1e-6) / 1e9
    for b in (l, r, o, ws):
        b.free()
    out = {
        "workload": "one rectified planar pair 1920x1080, 256 disparities (4..248 px), 8 paths, "
                    "Stage.Disparity's matcher alone",
        "ms_per_pair": round(ms, 4), "pairs_per_s": round(1e3 / ms, 2),
        "cost_evals_per_s": round(1e3 / ms * n, 1),
        "kernel_us": {k: round(v, 2) for k, v in us.items()},
        "roofline": {
            "aggregation": {"achieved": round(agg, 1), "frac": round(agg / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes": 6 * n + 8 * pixels},
            "census_cost": {"achieved": round(cost, 1), "frac": round(cost / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes": n + 2 * pixels},
            "peak": HBM_PEAK_GBS, "unit": "GB/s"},
    }
    traffic = load_traffic("1080")
    if traffic:
        t = traffic["aggregation_traffic_bytes"]
        out["roofline"]["aggregation"]["traffic"] = t
        out["roofline"]["aggregation"]["traffic_gbs"] = round(t / (aggregation_us * 1e-6) / 1e9, 1)
    return out


def _segment_rate(members, frames, runs):
    """Pairs/s of a plain pipeline segment at 728x544 on device-resident frames."""
    from nx_stereo_amd import hip, segment

    geometry = workload()[0]
    context = segment.Context(0)
    ok, seg = segment.HipSegment.build(members, geometry, context)
    reply = context.request(seg.configure_record())
    if reply is not True:
        raise RuntimeError(f"configure refused: {reply}")
    seg.configured_version = seg.version
    record = seg.run_record({})
    stream = hip.Stream()
    entries = [("left_frame", 0, (HEIGHT, WIDTH), frames[0].ptr),
               ("right_frame", 0, (HEIGHT, WIDTH), frames[1].ptr)]
    for _ in range(5):
        context.run_device(record, entries, stream=stream.handle)
    hip.synchronize()
    start, end = hip.Event(), hip.Event()
    start.record(stream)
    for _ in range(runs):
        context.run_device(record, entries, stream=stream.handle)
    end.record(stream)
    end.synchronize()
    ms = start.elapsed_ms(end) / runs
    context.close()
    return ms


def coarse_to_fine_leg(runs: int = 20) -> dict:
    """sgm_coarse_to_fine/3 (band 16, margin 4, 8 paths for both passes, sparse census)
    through the matcher alone at 728x544 and 1920x1080, and the plain pipeline with the
    Disparity stage's search coarse_to_fine at 728x544."""
    from nx_stereo_amd import _native, hip, segment, stages

    band, margin, paths = 16, 4, 8
    out = {"settings": {"band": band, "margin": margin, "paths": paths, "coarse_paths": paths,
                        "fine_census": "sparse_9x7", "disparities": 128}}
    for w, h in ((WIDTH, HEIGHT), (1920, 1080)):
        left, right = match_pair(w, h, 128, SEED + 5)
        l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
        ws = hip.Buffer(_native.coarse_to_fine_workspace_size(w, h, band, paths, paths))
        stream = hip.Stream()

        def match():
            _native.coarse_to_fine_device(l.ptr, r.ptr, w, h, band, margin, paths, paths, 0, 8, 64,
                                          1.0526315789473684, o.ptr, ws.ptr, ws.nbytes,
                                          stream.handle)

        for _ in range(3):
            match()
        stream.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(runs):
            match()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / runs
        segment.profile_start(16 * runs + 16)
        for _ in range(runs):
            match()
        stream.synchronize()
        kernels = segment.profile_stop()
        us = {k: v[0] / max(v[1], 1) * 1e3 for k, v in kernels.items() if v[1]}
        per_pair = {k: v[0] / runs * 1e3 for k, v in kernels.items() if v[1]}
        pixels, nb = w * h, w * h * band
        # The fine aggregation: band paths + band selection, the band volume's 6 Nb + 8P.
        agg_us = us["band_path"] + us["band_select"]
        agg = (6 * nb + 8 * pixels) / (agg_us * 1e-6) / 1e9
        out[f"{w}x{h}"] = {
            "workload": f"one rectified planar pair {w}x{h}, disparities 4..120, "
                        "sgm_coarse_to_fine/3 alone",
            "ms_per_pair": round(ms, 4), "pairs_per_s": round(1e3 / ms, 2),
            "kernel_us_per_launch": {k: round(v, 2) for k, v in us.items()},
            "kernel_us_per_pair": {k: round(v, 2) for k, v in per_pair.items()},
            "roofline": {"kernel": "band aggregation: band_path + band_select",
                         "achieved": round(agg, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(agg / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": 6 * nb + 8 * pixels},
        }
        for b in (l, r, o, ws):
            b.free()
        stream.destroy()
    geometry, members, left_frame, right_frame = workload()
    members = [(name, stages.Disparity(num_disparities=128, paths=paths, search="coarse_to_fine")
                if name == "disparity" else stage) for name, stage in members]
    frames = [hip.Buffer.of(f) for f in (left_frame, right_frame)]
    ms = _segment_rate(members, frames, runs)
    out["pipeline_728x544"] = {
        "workload": "the plain pipeline of the headline with Disparity search coarse_to_fine",
        "ms_per_pair": round(ms, 4), "pairs_per_s": round(1e3 / ms, 2)}
    for f in frames:
        f.free()
    return out


def streams_leg(pairs: int = 96) -> dict:
    """The main-line pipeline (and the same with search coarse_to_fine) as 1, 2 and 4
    segments, each on its own stream, frames resident, pairs enqueued round-robin: the
    throughput of several camera streams on one GPU. Each segment keeps its own carries."""
    from nx_stereo_amd import hip, segment, stages

    geometry, members, left_frame, right_frame = workload()
    c2f = [(name, stages.Disparity(num_disparities=128, paths=PATHS, search="coarse_to_fine")
            if name == "disparity" else stage) for name, stage in members]
    frames = [hip.Buffer.of(f) for f in (left_frame, right_frame)]
    entries = [("left_frame", 0, (HEIGHT, WIDTH), frames[0].ptr),
               ("right_frame", 0, (HEIGHT, WIDTH), frames[1].ptr)]
    out = {"workload": "the main-line pipeline as n segments on n streams, pairs round-robin, "
                       "frames resident; pairs/s over all segments"}
    for name, stage_list in (("full", members), ("coarse_to_fine", c2f)):
        rates = {}
        for n in (1, 2, 4):
            context = segment.Context(0)
            records = []
            for _ in range(n):
                ok, seg = segment.HipSegment.build(stage_list, geometry, context)
                if context.request(seg.configure_record()) is not True:
                    raise RuntimeError("configure refused")
                seg.configured_version = seg.version
                records.append(seg.run_record({}))
            streams = [hip.Stream() for _ in range(n)]

            def round_robin(count):
                for k in range(count):
                    context.run_device(records[k % n], entries, stream=streams[k % n].handle)

            round_robin(4 * n)
            hip.synchronize()
            start = time.perf_counter()
            round_robin(pairs)
            hip.synchronize()
            rates[str(n)] = round(pairs / (time.perf_counter() - start), 1)
            for st in streams:
                st.destroy()
            context.close()
        out[name] = rates
    for f in frames:
        f.free()
    return out


def host_io_leg(pairs: int = 200) -> dict:
    """The main-line pipeline as a caller with host buffers runs it: each pair's raw frames
    in a run record through nxs_submit (copied to pinned staging, uploaded), every output
    (rectified images, disparity, depth_mm, cloud, kept count, quality sums) downloaded into
    the reply. Pipelined (every pair submitted at once; the segment's three run slots keep
    up to three in flight), and one at a time (each reply waited for), as a Port's caller
    sees it."""
    from nx_stereo_amd import protocol, segment, serving

    geometry, members, left_frame, right_frame = workload()
    with segment.Context(0) as context:
        ok, seg = segment.HipSegment.build(members, geometry, context)
        if context.request(seg.configure_record()) is not True:
            raise RuntimeError("configure refused")
        seg.configured_version = seg.version
        record = seg.run_record({"left_frame": left_frame, "right_frame": right_frame})
        reply = context.request(record)
        if not isinstance(reply, protocol.Outputs):
            raise RuntimeError(f"run refused: {reply}")
        down = sum(t.nbytes for t in reply.tensors.values())
        arrivals = [[0.0] * pairs]
        serving.serve(context, [seg], [[record]], [[0.0] * 8], all_at_once=True)  # warm
        pipelined = serving.serve(context, [seg], [[record]], arrivals, all_at_once=True).summary()
        sequential = serving.serve_one_at_a_time(context, [seg], [[record]], [[0.0] * 50],
                                                 all_at_once=True).summary()
    return {"workload": "the main-line pipeline with 728x544 host frames in (792 KB a pair) and "
                        "every output back (%d bytes a pair), through nxs_submit" % down,
            "pipelined_pairs_per_s": pipelined["pairs_per_s"],
            "one_at_a_time_pairs_per_s": sequential["pairs_per_s"]}


BATCH_CLASSES = [(480, 270, 64, 24), (728, 544, 128, 48), (1280, 720, 192, 24)]


def batched_leg(replays: int = 5) -> dict:
    """96 pairs (24 x 480x270/64, 48 x 728x544/128, 24 x 1280x720/192) in one captured graph,
    each class's aggregation schedule autotuned on the batch, against the same 96 pairs as
    plain per-pair launches on one stream."""
    from nx_stereo_amd import _native, hip

    pairs, buffers = [], []
    for k, (w, h, d, count) in enumerate(BATCH_CLASSES):
        for j in range(count):
            left, right = match_pair(w, h, d, SEED + 1000 * (k + 1) + j)
            l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
            buffers += [l, r, o]
            pairs.append((l.ptr, r.ptr, o.ptr, w, h, d, 8, 8, 64, 1.0526315789473684))
    stream = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()

    def timed(batch):
        batch.run(stream.handle)
        stream.synchronize()
        e0.record(stream)
        for _ in range(replays):
            batch.run(stream.handle)
        e1.record(stream)
        e1.synchronize()
        describe = batch.describe()
        batch.close()
        return e0.elapsed_ms(e1) / replays, describe

    batched_ms, describe = timed(_native.Batch(pairs, autotune=True))
    # Every class in one launch per kernel, untuned (the tuned plans above include it).
    fused_ms, _ = timed(_native.Batch(pairs, autotune=False, fused=True))

    # Plain: one nxs_match_device per pair, one workspace per size reused pair after pair.
    spaces = {}
    for (_, _, _, w, h, d, paths, *_rest) in pairs:
        if (w, h, d) not in spaces:
            spaces[(w, h, d)] = hip.Buffer(_native.workspace_size(w, h, d, paths))

    def plain():
        for (lp, rp, op, w, h, d, paths, p1, p2, u) in pairs:
            ws = spaces[(w, h, d)]
            _native.match_device(lp, rp, w, h, d, paths, p1, p2, u, op, ws.ptr, ws.nbytes,
                                 stream.handle)

    plain()
    stream.synchronize()
    e0.record(stream)
    for _ in range(replays):
        plain()
    e1.record(stream)
    e1.synchronize()
    plain_ms = e0.elapsed_ms(e1) / replays
    for b in list(spaces.values()) + buffers:
        b.free()
    n = sum(w * h * d * c for w, h, d, c in BATCH_CLASSES)
    return {
        "workload": "96 rectified planar pairs: 24 x 480x270/64, 48 x 728x544/128, "
                    "24 x 1280x720/192, 8 paths, Stage.Disparity's matcher",
        "pairs": len(pairs), "cost_evals_per_batch": n,
        "batched_ms": round(batched_ms, 3), "batched_pairs_per_s": round(len(pairs) / batched_ms * 1e3, 2),
        "plain_ms": round(plain_ms, 3), "plain_pairs_per_s": round(len(pairs) / plain_ms * 1e3, 2),
        "speedup": round(plain_ms / batched_ms, 3),
        "fused_ms": round(fused_ms, 3), "fused_pairs_per_s": round(len(pairs) / fused_ms * 1e3, 2),
        "schedules": describe,
    }


SERVED_STREAMS = [(728, 544, 128, 30.0), (480, 270, 64, 60.0), (1280, 720, 192, 15.0)]


def served_leg(duration_s: float = 1.5) -> dict:
    """24 camera streams (8 x 728x544/128 at 30 Hz, 8 x 480x270/64 at 60 Hz, 8 x 1280x720/192
    at 15 Hz) with Poisson arrivals, each Disparity -> TemporalFilter -> Quality with its
    carries on the device, served continuously; then the same arrivals one pair at a time."""
    from nx_stereo_amd import segment, serving

    specs = [serving.StreamSpec(w, h, d, rate, SEED + 100 * k + j)
             for k, (w, h, d, rate) in enumerate(SERVED_STREAMS) for j in range(8)]
    arrivals = [serving.poisson_arrivals(spec.rate_hz, duration_s, spec.seed) for spec in specs]
    warm = 3
    frames = [serving.stream_frames(spec, len(a) + warm) for spec, a in zip(specs, arrivals)]
    result = {"workload": "24 streams with Poisson arrivals: 8 x 728x544/128 at 30 Hz, "
                          "8 x 480x270/64 at 60 Hz, 8 x 1280x720/192 at 15 Hz; planes drifting "
                          "0.5 px per frame; Disparity -> TemporalFilter -> Quality per stream, "
                          "host pairs in, disparity and quality sums back",
              "duration_s": duration_s,
              "offered_pairs_per_s": round(sum(len(a) for a in arrivals) / duration_s, 2)}
    with segment.Context(0) as ctx:
        segs = serving.configure_streams(ctx, specs)
        records = serving.run_records(segs, frames)
        warm_records = [r[:warm] for r in records]
        timed = [r[warm:] for r in records]
        serving.serve(ctx, segs, warm_records, [[0.0] * warm for _ in specs], all_at_once=True)
        before = ctx.coalescing()
        result["continuous"] = serving.serve(ctx, segs, timed, arrivals).summary()
        after = ctx.coalescing()
        result["continuous"]["coalesced"] = {"launches": after[0] - before[0],
                                             "pairs": after[1] - before[1]}
        # The same arrivals with every pair's matcher its own launch.
        ctx.coalescing(1)
        result["continuous_uncoalesced"] = serving.serve(ctx, segs, timed, arrivals).summary()
        ctx.coalescing(8)
        result["one_at_a_time"] = serving.serve_one_at_a_time(ctx, segs, timed, arrivals).summary()
        # Capacity: every pair of the trace submitted at once.
        result["saturated"] = {
            "continuous_pairs_per_s": serving.serve(ctx, segs, timed, arrivals,
                                                    all_at_once=True).summary()["pairs_per_s"],
            "one_at_a_time_pairs_per_s": serving.serve_one_at_a_time(
                ctx, segs, timed, arrivals, all_at_once=True).summary()["pairs_per_s"]}
        # Offered load against latency: the streams' rates scaled, 1 s of arrivals a point,
        # each stream's first 12 frames sent round again.
        sweep = serving.load_sweep(ctx, segs, [r[:12] for r in timed], specs, SWEEP_FACTORS,
                                   1.0, seed=1)
    result["load_sweep"] = sweep
    result["at_equal_p99"] = [serving.at_equal_p99(sweep, target) for target in SWEEP_P99_MS]
    return result


# The served leg's offered loads (multiples of the streams' 840 pairs/s) and the p99 targets
# at which the two forms' carried pairs/s are compared (the second is stream.exs's 33.3 ms
# frame budget at 30 Hz).
SWEEP_FACTORS = (0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0)
SWEEP_P99_MS = (5.0, 1000.0 / 30)


# ---------------------------------------------------------------------------------------
# HBM traffic from rocprofv3 counters. `--kernels-only SIZE` runs three matchings of the
# given workload and a 256 MiB device-to-device copy (the calibration) and nothing else, for
#   rocprofv3 --pmc FETCH_SIZE ... -- python bench.py --kernels-only 728
#   rocprofv3 --pmc WRITE_SIZE ... -- python bench.py --kernels-only 728
# and `--traffic FETCH_CSV WRITE_CSV OUT` turns the two counter_collection.csv files into
# OUT (profiles/traffic_*.json), which the default run reads into roofline.traffic.

TRAFFIC_FILES = {"728": os.path.join(ROOT, "profiles", "traffic_728x544x128.json"),
                 "1080": os.path.join(ROOT, "profiles", "traffic_1920x1080x256.json"),
                 "c2f728": os.path.join(ROOT, "profiles", "traffic_c2f_728x544.json"),
                 "c2f1080": os.path.join(ROOT, "profiles", "traffic_c2f_1920x1080.json")}
CALIBRATION_BYTES = 256 << 20


def kernels_only(size: str) -> int:
    from nx_stereo_amd import _native, hip

    if size in ("batch", "batch_fused"):
        # The batched leg's 96 pairs: autotuned (or every class in one launch per kernel),
        # then one run and five replays of the graph.
        pairs, buffers = [], []
        for k, (w, h, d, count) in enumerate(BATCH_CLASSES):
            for j in range(count):
                left, right = match_pair(w, h, d, SEED + 1000 * (k + 1) + j)
                l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
                buffers += [l, r, o]
                pairs.append((l.ptr, r.ptr, o.ptr, w, h, d, 8, 8, 64, 1.0526315789473684))
        batch = _native.Batch(pairs, autotune=size == "batch", fused=size == "batch_fused")
        stream = hip.Stream()
        for _ in range(6):
            batch.run(stream.handle)
        stream.synchronize()
        print(batch.describe().split("\n")[0], flush=True)
        batch.close()
        print(f"kernels-only {size} done", flush=True)
        return 0
    if size.startswith("c2f"):
        # sgm_coarse_to_fine/3 as the coarse_to_fine leg runs it (band 16, margin 4, 8 paths).
        w, h = (WIDTH, HEIGHT) if size == "c2f728" else (1920, 1080)
        left, right = match_pair(w, h, 128, SEED + 5)
        l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
        ws = hip.Buffer(_native.coarse_to_fine_workspace_size(w, h, 16, 8, 8))
        for _ in range(5):
            _native.coarse_to_fine_device(l.ptr, r.ptr, w, h, 16, 4, 8, 8, 0, 8, 64,
                                          1.0526315789473684, o.ptr, ws.ptr, ws.nbytes)
        hip.synchronize()
        print(f"kernels-only coarse-to-fine {w}x{h} done", flush=True)
        return 0
    # The kernels' accesses do not depend on the images' content: a planar pair of the size.
    if size == "728":
        w, h, d = WIDTH, HEIGHT, DISPARITIES
        left, right = match_pair(w, h, d, SEED)
    else:
        w, h, d = 1920, 1080, 256
        left, right = match_pair(w, h, d, SEED + 3)
    l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
    ws = hip.Buffer(_native.workspace_size(w, h, d, PATHS))
    for _ in range(3):
        _native.match_device(l.ptr, r.ptr, w, h, d, PATHS, 8, 64, 1.0526315789473684, o.ptr,
                             ws.ptr, ws.nbytes)
    a, b = hip.Buffer(CALIBRATION_BYTES), hip.Buffer(CALIBRATION_BYTES)
    hip.check(hip.runtime().hipMemcpy(ctypes_ptr(b.ptr), ctypes_ptr(a.ptr), CALIBRATION_BYTES, 3),
              "hipMemcpy")
    hip.synchronize()
    print(f"kernels-only {w}x{h}x{d} done", flush=True)
    return 0


def ctypes_ptr(address):
    import ctypes

    return ctypes.c_void_p(address)


def traffic_from(fetch_csv: str, write_csv: str, out: str) -> int:
    """Average FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of each matcher kernel, scaled
    by what the same counter reads for the 256 MiB copy (which reads and writes 256 MiB)."""
    import csv

    def per_kernel(path, counter):
        sums, counts = {}, {}
        for row in csv.DictReader(open(path)):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"].split("(")[0].split("::")[-1].split("<")[0]
            sums[name] = sums.get(name, 0.0) + float(row["Counter_Value"])
            counts[name] = counts.get(name, 0) + 1
        return {k: sums[k] / counts[k] for k in sums}

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise SystemExit("no calibration copy in the fetch counters")
    kib = CALIBRATION_BYTES / 1024
    fetch_factor, write_factor = kib / fetch[copy[0]], kib / write[copy[0]]
    kernels = {}
    for k in fetch:
        if "copyBuffer" in k:
            continue
        kernels[k] = {"fetch_bytes": round(fetch[k] * 1024 * fetch_factor),
                      "write_bytes": round(write.get(k, 0.0) * 1024 * write_factor)}
    aggregation = sum(v["fetch_bytes"] + v["write_bytes"] for k, v in kernels.items()
                      if k.startswith("path") or k == "select_kernel")
    result = {"fetch_factor": round(fetch_factor, 4), "write_factor": round(write_factor, 4),
              "calibration": "hipMemcpy device-to-device of 256 MiB (reads and writes 256 MiB)",
              "kernels": kernels, "aggregation_traffic_bytes": aggregation,
              "sources": [fetch_csv, write_csv]}
    with open(out, "w") as f:
        json.dump(result, f, indent=1)
        f.write("\n")
    print(json.dumps(result))
    return 0


def load_traffic(size: str):
    path = TRAFFIC_FILES[size]
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=200)
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    parser.add_argument("--legs", default="large,batched,served,coarse_to_fine,streams,host_io",
                        help="comma-separated extra legs: large, batched, served, "
                             "coarse_to_fine, streams, host_io (or none)")
    parser.add_argument("--kernels-only", choices=sorted(TRAFFIC_FILES) + ["batch", "batch_fused"],
                        help="run three matchings and the calibration copy (c2f*: five "
                             "coarse-to-fine matchings; batch: the batched leg's graph, "
                             "autotuned, then six runs), for rocprofv3")
    parser.add_argument("--traffic", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT"),
                        help="HBM bytes per launch from two rocprofv3 --pmc runs into OUT")
    args = parser.parse_args()
    if args.kernels_only:
        return kernels_only(args.kernels_only)
    if args.traffic:
        return traffic_from(*args.traffic)
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

    # The timed region: K steps back to back, an event on each side and nothing between
    # (an event or a profiling bracket between kernels costs several us of idle GPU).
    start, end = hip.Event(), hip.Event()
    hip.synchronize()
    wall = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    hip.synchronize()
    wall = time.perf_counter() - wall
    gpu_ms = start.elapsed_ms(end)
    # Per-pair GPU times (an event between pairs), then the kernels' times (HIP events around
    # every matcher kernel), each in a pass of its own after the timed one.
    events = [hip.Event() for _ in range(args.steps + 1)]
    events[0].record(stream)
    for k in range(args.steps):
        step()
        events[k + 1].record(stream)
    hip.synchronize()
    per_pair = np.array([events[k].elapsed_ms(events[k + 1]) for k in range(args.steps)])
    segment.profile_start(capacity=8 * args.steps + 16)
    for _ in range(args.steps):
        step()
    hip.synchronize()
    kernels = segment.profile_stop()

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
                       "note": "GPU time per pair, pairs enqueued back to back, an event "
                               "between pairs"},
        "wall_pairs_per_s": round(args.steps / wall, 2),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "aggregation: path_kernel + select_kernel",
            "algorithmic_bytes": algorithmic,
            "kernel_us": {k: round(v[0] / v[1] * 1e3, 2) for k, v in kernels.items() if v[1]},
        },
    }
    traffic = load_traffic("728")
    if traffic:
        result["roofline"]["traffic"] = traffic["aggregation_traffic_bytes"]
        result["roofline"]["traffic_note"] = (
            "rocprofv3 FETCH_SIZE x %.2f + WRITE_SIZE x %.2f of the path and select kernels, "
            "per pair (profiles/traffic_728x544x128.json); the working set fits the 256 MB "
            "Infinity Cache, so part of it never reaches HBM" % (traffic["fetch_factor"],
                                                                  traffic["write_factor"]))
    context.close()
    for frame in frames:
        frame.free()
    legs = {leg for leg in args.legs.split(",") if leg and leg != "none"}
    if "large" in legs:
        result["large_pair"] = large_pair_leg()
    if "batched" in legs:
        result["batched"] = batched_leg()
    if "served" in legs:
        result["served"] = served_leg()
    if "coarse_to_fine" in legs:
        result["coarse_to_fine"] = coarse_to_fine_leg()
    if "streams" in legs:
        result["streams"] = streams_leg()
    if "host_io" in legs:
        result["host_io"] = host_io_leg()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(geometry, members, left_frame, right_frame)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
