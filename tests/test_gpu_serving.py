"""The served form: streams of uneven sizes through one context at once.

Every stream's replies are the same, bit for bit, when its pairs are served in the
mix (all streams' pairs submitted together, overlapping on the device, admitted by
run slots and the HBM budget) as when the stream is served alone one pair at a time,
and both are the oracle's Disparity -> TemporalFilter -> Quality chain.
"""
import numpy as np
import pytest

from nx_stereo_amd import _native, protocol, segment, serving
from oracle import oracle

pytestmark = pytest.mark.gpu

# Two streams share a size class; the others differ in size and disparity range.
SPECS = [serving.StreamSpec(97, 31, 64, 60.0, 11), serving.StreamSpec(160, 48, 128, 30.0, 12),
         serving.StreamSpec(120, 40, 192, 15.0, 13), serving.StreamSpec(97, 31, 64, 60.0, 14),
         serving.StreamSpec(64, 48, 256, 30.0, 15)]
FRAMES = 5


def oracle_stream(spec, frames):
    temporal = oracle.Temporal((spec.height, spec.width))
    quality = oracle.Quality((spec.height, spec.width))
    out = []
    for left, right in frames:
        d = oracle.sgm(left, right, spec.disparities, 8)
        d = temporal.filter(d, 0.4, 2.0, 2)
        out.append((d, quality.sums(d)))
    return out


def decoded(reply):
    r = protocol.decode(reply)
    assert isinstance(r, protocol.Outputs), r
    return r.tensors["disparity"], r.tensors["quality_sums"]


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def frames():
    return [serving.stream_frames(spec, FRAMES) for spec in SPECS]


def alone(spec, fs):
    """The stream by itself in a fresh context, one pair at a time."""
    with segment.Context(0) as ctx:
        (seg,) = serving.configure_streams(ctx, [spec])
        out = []
        for left, right in fs:
            ticket = ctx.submit(seg.run_record({"left": left, "right": right}))
            assert ctx.wait(ticket, 60000)
            out.append(decoded(ctx.reply(ticket)))
        return out


@pytest.mark.parametrize("budget,coalesce,transfer", [(None, 8, "1"), (None, 1, "1"), (1, 8, "1"),
                                                      (None, 8, "0")])
def test_each_stream_is_the_same_alone_and_in_the_mix(frames, budget, coalesce, transfer,
                                                      monkeypatch):
    """budget=1 byte: every matcher run waits for the one before to free the device memory
    (and no run joins another's launch). coalesce=1: every run its own matcher launch.
    transfer="0": each segment's uploads and downloads on its own stream."""
    monkeypatch.setenv("NXS_TRANSFER_STREAMS", transfer)
    with segment.Context(0) as ctx:
        ctx.coalescing(coalesce)
        if budget is not None:
            assert ctx.lib.nxs_set_hbm_budget(ctx.handle, budget) == _native.NXS_OK
        segs = serving.configure_streams(ctx, SPECS)
        records = serving.run_records(segs, frames)
        arrivals = [np.zeros(FRAMES) for _ in SPECS]  # every pair at once: the most overlap
        served = serving.serve(ctx, segs, records, arrivals, keep_replies=True, all_at_once=True)
    mixed = [[None] * FRAMES for _ in SPECS]
    for s, k, reply in zip(served.stream, served.frame, served.replies):
        mixed[s][k] = decoded(reply)
    for s, spec in enumerate(SPECS):
        want = oracle_stream(spec, frames[s])
        solo = alone(spec, frames[s])
        for k in range(FRAMES):
            assert same(mixed[s][k][0], solo[k][0]) and same(mixed[s][k][1], solo[k][1]), (s, k)
            assert same(mixed[s][k][0], want[k][0]), ("disparity", s, k)
            assert same(mixed[s][k][1], want[k][1]), ("quality", s, k)


def test_poisson_arrivals_complete_in_order_per_stream(frames):
    """Arrivals at the streams' rates: every pair completes, and a stream's pairs complete in
    the order they arrived (its segment computes them in order)."""
    with segment.Context(0) as ctx:
        segs = serving.configure_streams(ctx, SPECS)
        records = serving.run_records(segs, frames)
        arrivals = [serving.poisson_arrivals(spec.rate_hz * 20, 10.0, spec.seed)[:FRAMES]
                    for spec in SPECS]
        served = serving.serve(ctx, segs, records, arrivals, keep_replies=True)
    for s in range(len(SPECS)):
        mine = served.stream == s
        order = np.argsort(served.arrival[mine], kind="stable")
        assert (np.diff(served.done[mine][order]) >= 0).all(), s
    assert (served.latency_ms > 0).all()
    summary = served.summary()
    assert summary["pairs"] == FRAMES * len(SPECS)


def test_a_segment_with_device_runs_does_not_block_admission():
    """With a 1-byte budget, a segment that has run through nxs_run_device (and keeps its
    workspace) must not hold back the admission of another segment's submitted run, and the
    context must close afterwards."""
    from nx_stereo_amd import hip

    specs = [serving.StreamSpec(64, 48, 64, 30.0, 1), serving.StreamSpec(97, 31, 64, 30.0, 2)]
    (left_a, right_a), = serving.stream_frames(specs[0], 1)
    (left_b, right_b), = serving.stream_frames(specs[1], 1)
    ctx = segment.Context(0)
    try:
        assert ctx.lib.nxs_set_hbm_budget(ctx.handle, 1) == _native.NXS_OK
        a, b = serving.configure_streams(ctx, specs)
        dl, dr = hip.Buffer.of(left_a), hip.Buffer.of(right_a)
        ctx.run_device(a.run_record({}), [("left", 0, (48, 64), dl.ptr),
                                          ("right", 0, (48, 64), dr.ptr)])
        hip.synchronize()
        ticket = ctx.submit(b.run_record({"left": left_b, "right": right_b}))
        assert ctx.wait(ticket, 60000), "the submitted run was never admitted"
        got, _ = decoded(ctx.reply(ticket))
        want = oracle.Temporal((31, 97)).filter(oracle.sgm(left_b, right_b, 64, 8), 0.4, 2.0, 2)
        assert same(got, want)
    finally:
        ctx.close()


def test_runs_on_a_caller_stream_and_submitted_runs_keep_their_order():
    """A temporal filter segment fed alternately through nxs_run_device on a caller's stream
    and through nxs_submit: each run sees the carry the run before it left, as when every
    run goes one way."""
    from nx_stereo_amd import hip, stages

    h, w = 16, 48
    ctx = segment.Context(0)
    stream = hip.Stream()
    try:
        geometry = {"width": w, "height": h, "frame_width": w, "frame_height": h}
        ok, seg = segment.HipSegment.build(
            [("temporal_filter", stages.TemporalFilter(alpha=0.5, delta=8.0, hold=0))],
            geometry, ctx)
        assert ok == "ok"
        assert ctx.request(seg.configure_record()) is True
        seg.configured_version = seg.version
        values = [8.0, 12.0, 6.0, 11.0, 14.0, 9.0]
        maps = [hip.Buffer.of(np.full((h, w), v, np.float32)) for v in values]
        got = []
        for k, v in enumerate(values):
            if k % 2 == 0:
                outs = ctx.run_device(seg.run_record({}), [("disparity", 2, (h, w), maps[k].ptr)],
                                      stream=stream.handle)
                stream.synchronize()
                got.append(float(hip.copy_to_host(outs[0][3], (h, w), np.float32)[0, 0]))
            else:
                ticket = ctx.submit(seg.run_record({"disparity": np.full((h, w), v, np.float32)}))
                assert ctx.wait(ticket, 60000)
                reply = protocol.decode(ctx.reply(ticket))
                got.append(float(reply.tensors["disparity"][0, 0]))
        want, previous = [], None
        for v in values:
            previous = v if previous is None else 0.5 * v + 0.5 * previous
            want.append(previous)
        assert got == want
    finally:
        stream.destroy()
        ctx.close()


# A 1280x720/192 stream whose fourth pair waits for one of its three run slots: the
# context's worker is held there while the pairs sent after it queue up.
HOLDER = serving.StreamSpec(1280, 720, 192, 15.0, 21)
# Three streams of one size class and one of another, interleaved in the queue.
GROUPED = [serving.StreamSpec(97, 31, 64, 60.0, 31), serving.StreamSpec(160, 48, 128, 30.0, 32),
           serving.StreamSpec(97, 31, 64, 60.0, 33), serving.StreamSpec(97, 31, 64, 60.0, 34)]


def submit_behind_a_held_worker(ctx, holder, records):
    """Four pairs of `holder`, then `records` (each (user, record)) in order; the holder's
    fourth pair waits for a run slot while the records queue. Returns the replies by user."""
    import ctypes
    import threading

    replies = {}
    finished = threading.Semaphore(0)

    @_native.DONE_FN
    def on_done(user, ticket, reply, length):
        replies[int(user or 0)] = ctypes.string_at(reply, length)
        finished.release()

    (left, right), = serving.stream_frames(HOLDER, 1)
    sends = [(1000 + k, holder.run_record({"left": left, "right": right})) for k in range(4)]
    sends += records
    buffers = [(user, (ctypes.c_uint8 * len(r)).from_buffer_copy(r)) for user, r in sends]
    for user, buf in buffers:
        assert ctx.lib.nxs_submit(ctx.handle, buf, len(buf), on_done, ctypes.c_void_p(user)) > 0
    for _ in buffers:
        assert finished.acquire(timeout=120), "a pair did not complete"
    return replies


def coalesced_within(attempt, attempts=5):
    """Runs `attempt` (a fresh context each time, every result checked; it returns the
    context's coalesced launches) until a launch was coalesced: the held worker makes that
    all but certain, and a retry covers a host slow enough to submit behind the worker."""
    for k in range(attempts):
        launches = attempt()
        if launches >= 1:
            return
        print(f"attempt {k}: nothing coalesced, again", flush=True)
    raise AssertionError(f"no launch coalesced in {attempts} attempts")


def test_queued_runs_of_one_size_class_share_a_matcher_launch():
    """The three 97x31/64 streams' pairs queued behind the held worker share matcher
    launches; every stream's replies are still the oracle's chain, in order."""
    frames = [serving.stream_frames(spec, 2) for spec in GROUPED]
    wants = [oracle_stream(spec, frames[s]) for s, spec in enumerate(GROUPED)]

    def attempt():
        with segment.Context(0) as ctx:
            ctx.coalescing(8)
            holder, *segs = serving.configure_streams(ctx, [HOLDER] + GROUPED)
            records = [(10 * s + k, segs[s].run_record({"left": l, "right": r}))
                       for k in range(2) for s, (l, r) in enumerate(f[k] for f in frames)]
            print("coalescing: 4 held pairs, then 8 pairs of two size classes", flush=True)
            replies = submit_behind_a_held_worker(ctx, holder, records)
            launches, pairs = ctx.coalescing()
        assert pairs >= 2 * launches, (launches, pairs)
        for s in range(len(GROUPED)):
            for k in range(2):
                got = decoded(replies[10 * s + k])
                assert same(got[0], wants[s][k][0]) and same(got[1], wants[s][k][1]), (s, k)
        return launches

    coalesced_within(attempt)


def test_a_refused_run_leaves_the_rest_of_its_group_computed():
    """One of the queued same-class pairs carries a penalty the matcher refuses: it alone is
    refused, with the stage's message; the others of its launch equal the oracle."""
    frames = [serving.stream_frames(spec, 1)[0] for spec in GROUPED]
    wants = {s: oracle_stream(GROUPED[s], [frames[s]])[0] for s in (0, 1, 3)}

    def attempt():
        with segment.Context(0) as ctx:
            ctx.coalescing(8)
            holder, *segs = serving.configure_streams(ctx, [HOLDER] + GROUPED)
            bad = segs[2]
            bad.inputs = [list(bad.inputs[0]), *bad.inputs[1:]]
            bad.inputs[0] = [(n, np.array([8.5], np.float32)) if n == "p1" else (n, v)
                             for n, v in bad.inputs[0]]
            records = [(s, seg.run_record({"left": l, "right": r}))
                       for s, (seg, (l, r)) in enumerate(zip(segs, frames))]
            print("coalescing: a refused pair among 4 queued pairs", flush=True)
            replies = submit_behind_a_held_worker(ctx, holder, records)
            launches, _ = ctx.coalescing()
        refused = protocol.decode(replies[2])
        assert isinstance(refused, protocol.Refused), refused
        assert refused.reason == "p1 must be an integer from 0 to 255, got 8.5"
        for s, want in wants.items():
            got = decoded(replies[s])
            assert same(got[0], want[0]) and same(got[1], want[1]), s
        return launches

    coalesced_within(attempt)


def pipeline_members():
    from nx_stereo_amd import stages

    return [("rectification", stages.Rectification()),
            ("disparity", stages.Disparity(num_disparities=64, paths=8)),
            ("speckle_filter", stages.SpeckleFilter()), ("depth", stages.Depth()),
            ("point_cloud", stages.PointCloud(decimation=4)), ("quality", stages.Quality())]


def test_coalesced_pipelines_with_stages_before_the_matcher_equal_each_alone():
    """Three Rectification -> Disparity -> SpeckleFilter -> Depth -> PointCloud -> Quality
    segments of one size, two pairs each, queued behind the held worker: their matchers share
    launches after each rectifies on its own stream, and every output of every pair (the
    second against the first's carried quality) equals the segment's pairs run alone."""
    from nx_stereo_amd import scenes
    from rig import rig_geometry

    geometry, cal = rig_geometry(182, 136)
    frames = [[scenes.raw_frames(cal, 40 + 2 * s + k, high=56.0)[:2] for k in range(2)]
              for s in range(3)]

    def segments_of(ctx, count):
        segs = []
        for _ in range(count):
            seg = segment.HipSegment(ctx, pipeline_members(), geometry)
            assert ctx.request(seg.configure_record()) is True
            seg.configured_version = seg.version
            segs.append(seg)
        return segs

    alone = {}
    for s in range(3):
        with segment.Context(0) as ctx:
            ctx.coalescing(1)
            (seg,) = segments_of(ctx, 1)
            for k, (lf, rf) in enumerate(frames[s]):
                ticket = ctx.submit(seg.run_record({"left_frame": lf, "right_frame": rf}))
                assert ctx.wait(ticket, 60000)
                alone[s, k] = protocol.decode(ctx.reply(ticket))

    def attempt():
        with segment.Context(0) as ctx:
            ctx.coalescing(8)
            (holder,) = serving.configure_streams(ctx, [HOLDER])
            segs = segments_of(ctx, 3)
            records = [(10 * s + k, segs[s].run_record({"left_frame": lf, "right_frame": rf}))
                       for k in range(2) for s, pairs in enumerate(frames)
                       for lf, rf in [pairs[k]]]
            print("coalescing: 3 pipelines x 2 pairs behind the held worker", flush=True)
            replies = submit_behind_a_held_worker(ctx, holder, records)
            launches, _ = ctx.coalescing()
        for (s, k), want_reply in alone.items():
            mixed = protocol.decode(replies[10 * s + k])
            assert isinstance(mixed, protocol.Outputs), mixed
            assert list(mixed.tensors) == list(want_reply.tensors), (s, k)
            for name, want in want_reply.tensors.items():
                got, want = np.asarray(mixed.tensors[name]), np.asarray(want)
                assert got.shape == want.shape and got.dtype == want.dtype, (s, k, name)
                assert got.tobytes() == want.tobytes(), (s, k, name)
        return launches

    coalesced_within(attempt)
