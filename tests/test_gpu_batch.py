"""The batched form (nxs_batch): pairs of several size classes in one captured graph,
every output bit for bit the oracle's, on every replay, with new images between replays;
and the autotuner's candidate schedules all give the same results."""
import numpy as np
import pytest

from nx_stereo_amd import _native, hip, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

# (width, height, disparities, paths, p1, p2, uniqueness)
CLASSES = [(97, 31, 64, 8, 8, 64, 1.0526316), (160, 48, 128, 8, 10, 70, 1.1),
           (120, 40, 192, 8, 8, 64, 1.0526316), (64, 48, 64, 3, 4, 40, 1.2)]


def make_pairs(seed):
    pairs = []
    for k, (w, h, d, paths, p1, p2, u) in enumerate(CLASSES):
        for j in range(2 if k != 1 else 3):
            left, right, _ = scenes.planar_pair(w, h, seed=seed * 100 + k * 10 + j, low=4,
                                                high=min(d - 8, 60))
            pairs.append((left, right, (w, h, d, paths, p1, p2, u)))
    return pairs


@pytest.mark.parametrize("fused", [False, True], ids=["tuned", "fused"])
def test_batch_matches_the_oracle_on_every_replay(fused):
    """Tuned: each class's schedule and the plan timed on the batch. Fused: the four classes
    (64 to 192 disparities, one at 3 paths) in one launch per kernel."""
    pairs = make_pairs(1)
    bufs = []
    for left, right, (w, h, d, paths, p1, p2, u) in pairs:
        bufs.append((hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)))
    batch = _native.Batch([(l.ptr, r.ptr, o.ptr, *cfg) for (l, r, o), (_, _, cfg) in zip(bufs, pairs)],
                          autotune=not fused, fused=fused)
    print(batch.describe(), flush=True)
    if fused:  # tuning may pick the fused plan as well
        assert batch.describe().startswith("plan: every class in one launch per kernel")
    stream = hip.Stream()
    for replay in range(3):
        if replay:
            pairs = [(*make_pairs(replay + 1)[i][:2], pairs[i][2]) for i in range(len(pairs))]
            for (left, right, _), (l, r, _o) in zip(pairs, bufs):
                l.upload(left)
                r.upload(right)
        batch.run(stream.handle)
        stream.synchronize()
        for (left, right, (w, h, d, paths, p1, p2, u)), (_l, _r, o) in zip(pairs, bufs):
            got = o.download((h, w), np.float32)
            want = oracle.sgm(left, right, d, paths, p1, p2, u)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (replay, w, h, d)
    batch.close()


# (branches, pairs a launch, fused checkpoints, work queue): the tuned plans, pinned one by
# one.
PLANS = [(1, 0, 0, 0), (1, 2, 0, 0), (2, 0, 0, 0), (2, 8, 0, 0), (4, 2, 0, 0), (1, 0, 4, 0),
         (1, 0, 8, 0), (2, 0, 4, 0), (2, 0, 8, 0), (4, 0, 8, 0), (1, 0, 8, 1), (1, 0, 4, 1),
         (2, 0, 8, 1)]


def test_every_plan_gives_the_oracle_results():
    """nxs_batch_create_planned: every plan the tuner times, pinned, on the mixed batch
    (four classes, 64 to 192 disparities, one at 3 paths), two replays each (the second
    replay of a work-queue plan starts from the counters the first left at zero)."""
    pairs = make_pairs(2)
    bufs = [(hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * cfg[0] * cfg[1]))
            for left, right, cfg in pairs]
    wants = [oracle.sgm(left, right, *cfg[2:]) for left, right, cfg in pairs]
    stream = hip.Stream()
    try:
        for plan in PLANS:
            print("plan", plan, flush=True)
            batch = _native.Batch([(l.ptr, r.ptr, o.ptr, *cfg)
                                   for (l, r, o), (_, _, cfg) in zip(bufs, pairs)], plan=plan)
            for replay in range(2):
                for _l, _r, o in bufs:
                    o.upload(np.zeros(o.nbytes // 4, np.float32))
                batch.run(stream.handle)
                stream.synchronize()
                for (left, right, (w, h, d, *_)), (_l, _r, o), want in zip(pairs, bufs, wants):
                    got = o.download((h, w), np.float32)
                    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (plan, replay, w, h, d)
            batch.close()
    finally:
        stream.destroy()
        for b in bufs:
            for x in b:
                x.free()


def test_a_fused_batch_of_eight_classes_up_to_256_disparities():
    """The most classes one fused launch holds: 64, 128, 192 and 256 disparities at 8 and 3
    paths (so the launches hold every disparity count's body, up to 16 a lane), two pairs
    a class, odd sizes, on two replays with new images between them."""
    classes = [(w, h, d, paths, 8, 64, 1.0526316)
               for (w, h), d in (((97, 31), 64), ((130, 37), 128), ((201, 29), 192),
                                 ((300, 41), 256))
               for paths in (8, 3)]
    assert len(classes) == 8

    def images(seed):
        return [scenes.planar_pair(w, h, seed=seed + 10 * k + j, low=4, high=min(d - 8, 80))[:2]
                for k, (w, h, d, *_r) in enumerate(classes) for j in range(2)]

    configs = [cfg for cfg in classes for _ in range(2)]
    pairs = images(500)
    bufs = [(hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * cfg[0] * cfg[1]))
            for (left, right), cfg in zip(pairs, configs)]
    batch = _native.Batch([(l.ptr, r.ptr, o.ptr, *cfg) for (l, r, o), cfg in zip(bufs, configs)],
                          autotune=False, fused=True)
    print(batch.describe(), flush=True)
    stream = hip.Stream()
    try:
        for replay in range(2):
            if replay:
                pairs = images(700)
                for (left, right), (l, r, _o) in zip(pairs, bufs):
                    l.upload(left)
                    r.upload(right)
            batch.run(stream.handle)
            stream.synchronize()
            for (left, right), cfg, (_l, _r, o) in zip(pairs, configs, bufs):
                w, h = cfg[:2]
                got = o.download((h, w), np.float32)
                want = oracle.sgm(left, right, *cfg[2:])
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (replay, cfg)
    finally:
        batch.close()
        stream.destroy()
        for b in bufs:
            for x in b:
                x.free()


def test_a_fused_batch_of_extreme_shapes():
    """Classes whose blocks a pair differ by hundreds of times in each fused launch: 4096-wide
    rows, 700-tall columns, a 9x7 pair (lines shorter than a checkpoint segment), 3 paths
    (eight down-only columns a block), one to three pairs a class."""
    classes = [((4096, 8, 64, 8), 1), ((16, 700, 64, 8), 2), ((9, 7, 64, 3), 3),
               ((33, 17, 128, 8), 2), ((250, 60, 192, 3), 1), ((70, 90, 256, 8), 2)]
    configs = [(w, h, d, paths, 8, 64, 1.0526316) for (w, h, d, paths), n in classes
               for _ in range(n)]
    pairs = [scenes.planar_pair(w, h, seed=900 + k, low=1, high=min(d - 8, 40, max(w - 2, 2)))[:2]
             for k, (w, h, d, *_r) in enumerate(configs)]
    bufs = [(hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * cfg[0] * cfg[1]))
            for (left, right), cfg in zip(pairs, configs)]
    batch = _native.Batch([(l.ptr, r.ptr, o.ptr, *cfg) for (l, r, o), cfg in zip(bufs, configs)],
                          autotune=False, fused=True)
    print(batch.describe(), flush=True)
    stream = hip.Stream()
    try:
        batch.run(stream.handle)
        stream.synchronize()
        for (left, right), cfg, (_l, _r, o) in zip(pairs, configs, bufs):
            w, h = cfg[:2]
            got = o.download((h, w), np.float32)
            want = oracle.sgm(left, right, *cfg[2:])
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), cfg
    finally:
        batch.close()
        stream.destroy()
        for b in bufs:
            for x in b:
                x.free()


LANES = {64: (16, 32), 128: (16, 32, 64), 192: (16,), 256: (16, 32, 64)}
WAVES = {64: (2, 4), 128: (2, 4, 8), 192: (2,), 256: (2, 4, 8)}  # more than one wave per line


@pytest.mark.parametrize("disparities,paths", [(64, 8), (128, 8), (64, 3), (128, 3), (192, 8),
                                               (256, 8), (256, 3)])
def test_every_schedule_gives_the_same_disparity(disparities, paths):
    """Tuning picks by time only: each candidate schedule is checked bit for bit here, on
    the path sums themselves as well as the output."""
    w, h = 97, 31
    left, right = scenes.scene(w, h, 7)
    want = oracle.sgm(left, right, disparities, paths)
    want_sums = oracle.aggregate(oracle.costs(oracle.census(left), oracle.census(right), disparities),
                                 paths, 8, 64)
    l, r, o = hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)
    ws = hip.Buffer(_native.workspace_size(w, h, disparities, paths))
    kept = _native.get_schedule(w, h, disparities, paths)
    try:
        two_lanes = (16, 32) if disparities in (64, 128, 256) else (16,)
        shapes = [(lanes, 1, 0) for lanes in LANES[disparities]] + \
            [(16, w, 0) for w in WAVES[disparities] if paths == 8 or w == 2] + \
            ([(32, 2, 0)] if disparities in (64, 128, 256) else []) + \
            [(lanes, 2, 1) for lanes in two_lanes]  # meeting by recomputation
        for lanes, waves, meet in shapes:
            for prefetch in (2, 4, 8):
                for split in (False, True):
                    _native.set_schedule(w, h, disparities, paths, prefetch, split, lanes, waves,
                                         meet)
                    _native.match_device(l.ptr, r.ptr, w, h, disparities, paths, 8, 64, 1.0526316,
                                         o.ptr, ws.ptr, ws.nbytes)
                    hip.synchronize()
                    got = o.download((h, w), np.float32)
                    case = (lanes, waves, meet, prefetch, split)
                    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), case
                    sums = _native.match_host(left, right, disparities, paths, intermediates=True)["sums"]
                    assert np.array_equal(sums, want_sums), case
    finally:
        _native.set_schedule(w, h, disparities, paths, *kept)


@pytest.mark.parametrize("disparities,paths", [(64, 8), (128, 8), (64, 3), (128, 3)])
def test_several_waves_on_odd_and_short_lines(disparities, paths):
    """Two or more waves per line (16 lanes a line, and two waves of 32): lines shorter than
    the prefetch and LDS windows, odd and even lengths (the middle pixel), every diagonal
    length from 1 up; at 3 paths, column counts off the eight (or four) a block takes."""
    sizes = [(9, 7), (10, 8), (17, 40), (33, 17)]
    kept = {size: _native.get_schedule(*size, disparities, paths) for size in sizes}
    try:
        for w, h in sizes:
            left, right = scenes.scene(w, h, 3)
            want = oracle.aggregate(oracle.costs(oracle.census(left), oracle.census(right),
                                                 disparities), paths, 8, 64)
            shapes = [(16, waves, 0) for waves in (WAVES[disparities] if paths == 8 else (2,))]
            for lanes, waves, meet in shapes + [(32, 2, 0), (16, 2, 1), (32, 2, 1)]:
                for prefetch in (2, 4, 8):
                    _native.set_schedule(w, h, disparities, paths, prefetch, False, lanes, waves,
                                         meet)
                    sums = _native.match_host(left, right, disparities, paths,
                                              intermediates=True)["sums"]
                    assert np.array_equal(sums, want), (w, h, lanes, waves, meet, prefetch)
    finally:
        for size, schedule in kept.items():
            _native.set_schedule(*size, disparities, paths, *schedule)


def test_lanes_that_do_not_fit_are_refused():
    with pytest.raises(_native.NativeError, match="do not fit"):
        _native.set_schedule(97, 31, 192, 8, 4, False, 32)
    with pytest.raises(_native.NativeError, match="do not fit"):
        _native.set_schedule(97, 31, 64, 8, 4, False, 64)
    with pytest.raises(_native.NativeError, match="16 lanes per line, or 32"):
        _native.set_schedule(97, 31, 128, 8, 4, False, 64, 2)
    with pytest.raises(_native.NativeError, match="four or more waves per line take 16 lanes"):
        _native.set_schedule(97, 31, 128, 8, 4, False, 32, 4)
    with pytest.raises(_native.NativeError, match="8 waves per line do not fit"):
        _native.set_schedule(97, 31, 64, 8, 4, False, 16, 8)
    with pytest.raises(_native.NativeError, match="4 waves per line take 8 paths"):
        _native.set_schedule(97, 31, 128, 3, 4, False, 16, 4)
    with pytest.raises(_native.NativeError, match="recomputation takes two waves per line"):
        _native.set_schedule(97, 31, 128, 8, 4, False, 16, 1, 1)
    with pytest.raises(_native.NativeError, match="meet must be 0 or 1"):
        _native.set_schedule(97, 31, 128, 8, 4, False, 16, 2, 2)


@pytest.mark.parametrize("w,h,disparities,paths", [
    (4096, 8, 64, 8),      # the widest rows: 512 segments of 8 a line
    (16, 700, 64, 8),      # columns and diagonals far longer than the rows
    (300, 40, 192, 8),     # 12 bytes a lane
    (4096, 8, 128, 3),     # rows meet, columns go down only
    (500, 120, 256, 8),
])
def test_meeting_by_recomputation_at_the_extremes(w, h, disparities, paths):
    """The two waves meeting by recomputation on very long and very short lines, every
    segment length, 16 and 32 lanes a line where they fit: the path sums bit for bit."""
    left, right, _ = scenes.planar_pair(w, h, seed=w + h, low=4, high=min(disparities - 8, 60))
    want = oracle.aggregate(oracle.costs(oracle.census(left), oracle.census(right), disparities),
                            paths, 8, 64)
    kept = _native.get_schedule(w, h, disparities, paths)
    lanes = (16, 32) if disparities in (64, 128, 256) else (16,)
    try:
        for lane_count in lanes:
            for prefetch in (2, 8):
                _native.set_schedule(w, h, disparities, paths, prefetch, False, lane_count, 2, 1)
                sums = _native.match_host(left, right, disparities, paths, intermediates=True)["sums"]
                assert np.array_equal(sums, want), (w, h, disparities, paths, lane_count, prefetch)
    finally:
        _native.set_schedule(w, h, disparities, paths, *kept)


def test_many_short_pairs_in_one_batch():
    """Two classes of many pairs only 7 or 8 rows tall (the least the matcher takes): rows
    clamped at both edges of every pair, and census launches of 630 and 640 rows over 90 and
    80 pairs: every pair's disparity bit for bit."""
    pairs, bufs = [], []
    for k, (w, h) in enumerate([(41, 7)] * 90 + [(33, 8)] * 80):
        left, right = scenes.scene(w, h, 100 + k)
        pairs.append((left, right))
        bufs.append((hip.Buffer.of(left), hip.Buffer.of(right), hip.Buffer(4 * w * h)))
    batch = _native.Batch([(l.ptr, r.ptr, o.ptr, left.shape[1], left.shape[0], 64, 8, 8, 64, 1.0526316)
                           for (l, r, o), (left, _) in zip(bufs, pairs)], autotune=False)
    stream = hip.Stream()
    batch.run(stream.handle)
    stream.synchronize()
    for (left, right), (_l, _r, o) in zip(pairs, bufs):
        got = o.download(left.shape, np.float32)
        want = oracle.sgm(left, right, 64, 8, 8, 64, 1.0526316)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), left.shape
    batch.close()
