"""The benchmark's workloads at their own sizes on the GPU, every output against the CPU
checker: the 1920x1080 pair at 256 disparities, the 96-pair batch of three size classes,
and the 24-stream served mix. The checker is the multi-threaded stereo_match-structured
matcher (oracle.cpu_sgm), which tests/test_oracle.py pins equal to the oracle bit for bit
at 64-256 disparities (the oracle's full i32 volumes do not fit these sizes), and the
oracle's temporal filter and quality sums."""
import numpy as np
import pytest

from nx_stereo_amd import _native, hip, protocol, scenes, segment, serving
from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 20261203
UNIQUENESS = 1.0526315789473684


def same_bits(got, want, what):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    bad = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{what}: {len(bad)} pixels differ, first at {tuple(bad[0])}"


@pytest.mark.parametrize("paths", [8, 3])
def test_large_pair_1920x1080x256(paths):
    """The large-pair workload (bench.py's large_pair leg: seed + 3, disparities 4..248)."""
    w, h, d = 1920, 1080, 256
    left, right, _ = scenes.planar_pair(w, h, seed=SEED + 3, low=4, high=d - 8)
    print(f"large pair {w}x{h}x{d}, {paths} paths on cuda:0", flush=True)
    l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
    ws = hip.Buffer(_native.workspace_size(w, h, d, paths))
    try:
        for _ in range(2):  # the second run reuses the workspace the first left behind
            _native.match_device(l.ptr, r.ptr, w, h, d, paths, 8, 64, UNIQUENESS, o.ptr, ws.ptr,
                                 ws.nbytes)
            hip.synchronize()
            got = o.download((h, w), np.float32)
            same_bits(got, oracle.cpu_sgm(left, right, d, paths), f"{w}x{h}x{d} paths {paths}")
        assert (got > 0).mean() > 0.5, "the planar scene is mostly matched"
    finally:
        for b in (l, r, o, ws):
            b.free()


# (width, height, disparities, pairs): bench.py's batched leg, 1:2:1.
BATCH_CLASSES = [(480, 270, 64, 24), (728, 544, 128, 48), (1280, 720, 192, 24)]


def batch_images(variant):
    """96 distinct pairs in the classes' 1:2:1 mix, interleaved (0, 1, 1, 2, ...): one planar
    pair per class, each pair of it rolled by its own number of rows (the rows' disparities
    roll with them) and, for the second variant, by a further 5."""
    base = {c: scenes.planar_pair(w, h, seed=SEED + 1000 * (c + 1), low=4, high=d - 8)[:2]
            for c, (w, h, d, _) in enumerate(BATCH_CLASSES)}
    order = [0, 1, 1, 2] * 24
    seen = [0, 0, 0]
    pairs = []
    for c in order:
        w, h, d, _ = BATCH_CLASSES[c]
        shift = (seen[c] * 37 + 5 * variant) % h
        seen[c] += 1
        left, right = (np.ascontiguousarray(np.roll(x, shift, 0)) for x in base[c])
        pairs.append((left, right, (w, h, d)))
    return pairs


@pytest.mark.parametrize("form", ["per_class", "fused", "fused_queue"])
def test_batch_of_96_pairs_at_the_benchmark_classes(form):
    """nxs_batch over the batched leg's classes, every pair checked after a replay, then
    again with new images uploaded before the next replay: launches per class, every class
    in one launch per kernel, and that with the aggregation's persistent blocks taking their
    line groups from a work queue."""
    pairs = batch_images(0)
    bufs = [(hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h))
            for left, right, (w, h, d) in pairs]
    print(f"batch: 96 pairs, 24 x 480x270/64, 48 x 728x544/128, 24 x 1280x720/192, {form}",
          flush=True)
    batch = _native.Batch([(l.ptr, r.ptr, o.ptr, w, h, d, 8, 8, 64, UNIQUENESS)
                           for (l, r, o), (_, _, (w, h, d)) in zip(bufs, pairs)], autotune=False,
                          fused=form == "fused",
                          plan=(1, 0, 8, 1) if form == "fused_queue" else None)
    stream = hip.Stream()
    try:
        for variant in range(2):
            if variant:
                pairs = batch_images(variant)
                for (left, right, _), (l, r, _o) in zip(pairs, bufs):
                    l.upload(left)
                    r.upload(right)
            batch.run(stream.handle)
            stream.synchronize()
            for k, ((left, right, (w, h, d)), (_l, _r, o)) in enumerate(zip(pairs, bufs)):
                got = o.download((h, w), np.float32)
                same_bits(got, oracle.cpu_sgm(left, right, d, 8), f"variant {variant} pair {k}")
    finally:
        batch.close()
        stream.destroy()
        for b in bufs:
            for x in b:
                x.free()


# bench.py's served leg: 8 streams each at 728x544/128 (30 Hz), 480x270/64 (60 Hz) and
# 1280x720/192 (15 Hz).
SERVED = [(728, 544, 128, 30.0), (480, 270, 64, 60.0), (1280, 720, 192, 15.0)]
FRAMES = 3


def served_specs():
    specs = []
    for k, (w, h, d, hz) in enumerate(SERVED):
        for j in range(8):
            specs.append(serving.StreamSpec(w, h, d, hz, SEED + 100 * k + j))
    return specs


def decoded(reply):
    r = protocol.decode(reply)
    assert isinstance(r, protocol.Outputs), r
    return r.tensors["disparity"], r.tensors["quality_sums"]


def test_24_served_streams_are_the_same_alone_in_the_mix_and_on_the_cpu():
    """Every stream's pairs (Disparity -> TemporalFilter -> Quality, carries on the device)
    submitted all at once with the others', against the stream served alone one pair at a
    time in a context of its own, and against the CPU chain."""
    specs = served_specs()
    frames = [serving.stream_frames(spec, FRAMES) for spec in specs]
    print(f"served: {len(specs)} streams x {FRAMES} frames, all at once", flush=True)
    with segment.Context(0) as ctx:
        segs = serving.configure_streams(ctx, specs)
        records = serving.run_records(segs, frames)
        arrivals = [np.zeros(FRAMES) for _ in specs]
        served = serving.serve(ctx, segs, records, arrivals, keep_replies=True, all_at_once=True)
        launches, pairs = ctx.coalescing()
        print(f"coalesced: {pairs} pairs in {launches} matcher launches", flush=True)
    mixed = [[None] * FRAMES for _ in specs]
    for s, k, reply in zip(served.stream, served.frame, served.replies):
        mixed[s][k] = decoded(reply)
    for s, spec in enumerate(specs):
        print(f"stream {s}: {spec.width}x{spec.height}/{spec.disparities} alone", flush=True)
        with segment.Context(0) as ctx:
            (seg,) = serving.configure_streams(ctx, [spec])
            alone = []
            for left, right in frames[s]:
                ticket = ctx.submit(seg.run_record({"left": left, "right": right}))
                assert ctx.wait(ticket, 120000)
                alone.append(decoded(ctx.reply(ticket)))
        temporal = oracle.Temporal((spec.height, spec.width))
        quality = oracle.Quality((spec.height, spec.width))
        for k, (left, right) in enumerate(frames[s]):
            d = temporal.filter(oracle.cpu_sgm(left, right, spec.disparities, 8), 0.4, 2.0, 2)
            q = quality.sums(d)
            same_bits(mixed[s][k][0], alone[k][0], f"stream {s} frame {k}: mix vs alone")
            same_bits(mixed[s][k][1], alone[k][1], f"stream {s} frame {k}: quality, mix vs alone")
            same_bits(mixed[s][k][0], d, f"stream {s} frame {k}: disparity vs cpu")
            same_bits(mixed[s][k][1], q, f"stream {s} frame {k}: quality vs cpu")
