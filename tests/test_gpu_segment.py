"""HipSegment through the C ABI on the GPU, against the oracle: every stage alone
and the full pipeline, bit for bit (the quality sums within 1e-5 relative, the
reference's own bar, native_parity_test.exs:102-110), and the protocol's
refusals, ordering and carries (segment.rs:574-778; main.rs:188-221;
stage/stateful_pipeline_test.exs:59-120)."""
import numpy as np
import pytest

from nx_stereo_amd import protocol, scenes, stages
from nx_stereo_amd.segment import Context, HipSegment
from oracle import oracle
from rig import rig_geometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def context():
    with Context(0) as ctx:
        yield ctx


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (a.shape, b.shape, a.dtype, b.dtype)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{len(bad)} differ, first at {tuple(bad[0])}"


def build(context, members, geometry):
    ok, seg = HipSegment.build(members, geometry, context)
    assert ok == "ok", seg
    return seg


def run_ok(seg, entries):
    result = seg.run(entries)
    assert result[0] == "ok", result
    return result[1]


def test_rectification_with_distortion_and_shrink(context):
    geometry, cal = rig_geometry(182, 136, 364, 272)
    frames = [scenes.noise(364, 272, 1), scenes.noise(364, 272, 2)]
    seg = build(context, [("rectification", stages.Rectification(right_row_offset=1.5))], geometry)
    out = run_ok(seg, {"left_frame": frames[0], "right_frame": frames[1]})
    params = seg.members[0][1].params(geometry)
    for side, frame in zip(("left", "right"), frames):
        cam = params[side]
        want = oracle.rectify(frame, cam["k"], cam["d"], cam["inverse_pr"], (136, 182))
        same_bits(out[side], want)


@pytest.mark.parametrize("disparities,paths", [(64, 3), (64, 8), (128, 8), (192, 8), (256, 8)])
def test_disparity_stage(context, disparities, paths):
    left, right, _ = scenes.planar_pair(320, 96, seed=disparities + paths, low=4,
                                        high=min(disparities - 8, 120))
    geometry, _ = rig_geometry(320, 96)
    stage = stages.Disparity(num_disparities=disparities, paths=paths, p1=10, p2=70, uniqueness=1.1)
    seg = build(context, [("disparity", stage)], geometry)
    out = run_ok(seg, {"left": left, "right": right})
    same_bits(out["disparity"], oracle.sgm(left, right, disparities, paths, 10, 70, 1.1))
    assert seg.member_times()[0][0] == "disparity"


def test_speckle_depth_cloud_quality_chain(context):
    geometry, _ = rig_geometry(200, 120)
    rng = np.random.default_rng(4)
    disparity = np.where(rng.random((120, 200)) < 0.2, -1.0,
                         rng.uniform(0.2, 60.0, (120, 200))).astype(np.float32)
    disparity[30:60, 40:90] = 17.25
    members = [("speckle_filter", stages.SpeckleFilter(min_support=10, max_diff=1.5)),
               ("depth", stages.Depth()),
               ("point_cloud", stages.PointCloud(decimation=3)),
               ("quality", stages.Quality())]
    seg = build(context, members, geometry)
    quality = oracle.Quality((120, 200))
    for pair in range(3):
        entry = np.roll(disparity, pair * 7, axis=1)
        out = run_ok(seg, {"disparity": entry})
        assert list(out) == ["disparity", "depth_mm", "cloud_grid", "cloud_points", "quality_sums"]
        d = oracle.despeckle(entry, 10, 1.5)
        same_bits(out["disparity"], d)
        mm, m = oracle.depth(d, geometry["focal_length_px"], geometry["baseline_m"])
        same_bits(out["depth_mm"], mm)
        cx, cy = geometry["principal_point"]
        points, kept = oracle.cloud(m, 3, geometry["focal_length_px"], cx, cy)
        same_bits(out["cloud_grid"], points)
        assert int(out["cloud_points"]) == kept
        want = quality.sums(d)
        np.testing.assert_allclose(out["quality_sums"], want, rtol=1e-5)


@pytest.mark.parametrize("radius", [1, 2, 3, 5])
def test_speckle_radius(context, radius):
    """Radius is a build constant (3 from Stage.SpeckleFilter; segment.rs takes any): the
    LDS-tiled kernel serves 3, the direct one the others. Sizes off the 32x8 tiles."""
    h, w = 45, 77
    rng = np.random.default_rng(radius)
    entry = np.where(rng.random((h, w)) < 0.25, -1.0, rng.uniform(0.5, 9.0, (h, w))).astype(np.float32)
    g = {"width": w, "height": h, "frame_width": w, "frame_height": h}
    segment_id = 7100 + radius
    settings = [("radius", float(radius))]
    assert context.request(protocol.configure(segment_id, g, [("speckle_filter", settings)])) is True
    inputs = [[("min_support", np.float32(5)), ("max_diff", np.float32(1.0))]]
    reply = context.request(protocol.run(segment_id, inputs, [("disparity", entry)]))
    same_bits(reply.tensors["disparity"], oracle.despeckle(entry, 5, 1.0, radius))
    context.request(protocol.release(segment_id))


@pytest.mark.parametrize("kernel_size", [3, 5])
def test_median_filter(context, kernel_size):
    """filters.rs:134-216 through stage/median_filter.ex's native kind, on a speckled map and
    on filters.rs:331-357's image (negatives and zeros). Sizes off the 256-wide rows."""
    h, w = 37, 301
    rng = np.random.default_rng(kernel_size)
    speckled = np.where(rng.random((h, w)) < 0.3, -1.0, rng.uniform(0.5, 60.0, (h, w)))
    formula = ((np.arange(h * w) * 37 + 11) % 53 / 4.0 - 1.0).reshape(h, w)
    geometry, _ = rig_geometry(w, h)
    seg = build(context, [("median_filter", stages.MedianFilter(kernel_size=kernel_size))], geometry)
    for entry in (speckled.astype(np.float32), formula.astype(np.float32)):
        out = run_ok(seg, {"disparity": entry})
        same_bits(out["disparity"], oracle.median(entry, kernel_size))


def test_sharpness(context):
    """measures.rs:11-51 through stage/sharpness.ex's native kind: a random pair with black
    pixels, and diagnostics_test.exs:17-25's closed form."""
    rng = np.random.default_rng(12)
    h, w = 70, 333
    left = np.where(rng.random((h, w)) < 0.1, 0, rng.integers(1, 256, (h, w))).astype(np.uint8)
    right = np.where(rng.random((h, w)) < 0.1, 0, rng.integers(1, 256, (h, w))).astype(np.uint8)
    geometry, _ = rig_geometry(w, h)
    seg = build(context, [("sharpness", stages.Sharpness())], geometry)
    for _ in range(2):  # the totals are set back to zero for the next pair
        out = run_ok(seg, {"left": left, "right": right})
        same_bits(out["sharpness_sums"], oracle.sharpness_sums(left, right))
    small = {"width": 3, "height": 2, "frame_width": 3, "frame_height": 2}
    assert context.request(protocol.configure(7300, small, [("sharpness", [])])) is True
    reply = context.request(protocol.run(7300, [[]], [
        ("left", np.array([[10, 20, 10], [10, 10, 10]], np.uint8)),
        ("right", np.full((2, 3), 10, np.uint8))]))
    assert reply.tensors["sharpness_sums"].tolist() == [200.0, 0.0, 50.0, 40.0, 4.0, 4.0]
    context.request(protocol.release(7300))


@pytest.mark.parametrize("bands, reach, row_step", [(4, 2, 4), (3, 9, 1), (7, 0, 5)])
def test_row_alignment(context, bands, reach, row_step):
    """measures.rs:53-108 through stage/row_alignment.ex's native kind (its constants 4, 2,
    4) and other build settings; reach 9 sums its 19 row offsets in two passes."""
    rng = np.random.default_rng(bands * 100 + reach)
    h, w = 61, 290
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    right = rng.integers(0, 256, (h, w), dtype=np.uint8)
    disparity = np.where(rng.random((h, w)) < 0.3, -1.0, rng.uniform(0.0, 80.0, (h, w)))
    disparity[::7, ::5] = np.floor(disparity[::7, ::5]) + 0.5  # halves round away from zero
    disparity = disparity.astype(np.float32)
    g = {"width": w, "height": h, "frame_width": w, "frame_height": h}
    settings = [("bands", float(bands)), ("reach", float(reach)), ("row_step", float(row_step))]
    segment_id = 7400 + bands
    assert context.request(protocol.configure(segment_id, g, [("row_alignment", settings)])) is True
    for _ in range(2):
        reply = context.request(protocol.run(segment_id, [[]], [
            ("left", left), ("right", right), ("disparity", disparity)]))
        got = reply.tensors["alignment_sums"]
        assert got.shape == (bands, 2 * reach + 2)
        same_bits(got, oracle.alignment_sums(left, right, disparity, bands, reach, row_step))
    context.request(protocol.release(segment_id))


def test_new_kinds_refuse_settings_as_segment_rs_does(context):
    g = {"width": 32, "height": 24, "frame_width": 32, "frame_height": 24}
    reply = context.request(protocol.configure(7500, g, [("median_filter", [("kernel_size", 4.0)])]))
    assert reply.reason == "median_filter kernel_size must be one of [3, 5], got 4"
    reply = context.request(protocol.configure(7500, g, [("row_alignment", [
        ("bands", 0.0), ("reach", 2.0), ("row_step", 4.0)])]))
    assert reply.reason == "row_alignment bands must be an integer from 1 to 32, got 0"
    reply = context.request(protocol.configure(7500, g, [("sharpness", [("radius", 3.0)])]))
    assert reply.reason == "sharpness has no setting radius"


C2F_STAGE_CASES = [
    {},
    dict(coarse_paths=3, band=32, margin=8, fine_census="dense_9x5_blend"),
    dict(paths=3, band=8, margin=0, fine_census="dense_9x5"),
]


@pytest.mark.parametrize("opts", C2F_STAGE_CASES)
def test_disparity_stage_coarse_to_fine(context, opts):
    # stage/disparity.ex:149-162: sgm_coarse_to_fine/3, then "> 0 else -1".
    left, right, _ = scenes.planar_pair(320, 96, seed=5, low=4, high=110)
    geometry, _ = rig_geometry(320, 96)
    stage = stages.Disparity(num_disparities=128, search="coarse_to_fine", p1=10, p2=70,
                             uniqueness=1.1, **opts)
    seg = build(context, [("disparity", stage)], geometry)
    out = run_ok(seg, {"left": left, "right": right})
    want = oracle.coarse_to_fine(left, right, band=stage.state["band"],
                                 margin=stage.state["margin"], paths=stage.state["paths"],
                                 coarse_paths=stage.state["coarse_paths"],
                                 fine_census=stage.state["fine_census"], p1=10, p2=70,
                                 uniqueness=1.1, positive=True)
    same_bits(out["disparity"], want)
    assert (want > 0).mean() > 0.5


def test_coarse_to_fine_settings_refused_as_the_stage_does(context):
    g = {"width": 322, "height": 96, "frame_width": 322, "frame_height": 96}
    base = [("num_disparities", 128.0), ("paths", 8.0)]
    cases = [
        (base + [("search", 1.0)], None),
        ([("num_disparities", 64.0), ("paths", 8.0), ("search", 1.0)],
         "coarse_to_fine searches 128 disparities, got 64"),
        (base + [("search", 1.0), ("band", 16.0), ("margin", 16.0)],
         "margin must be below band (16), got 16"),
        (base + [("search", 1.0), ("band", 12.0)],
         "disparity band must be one of [8, 16, 32, 64], got 12"),
        (base + [("fine_census", 1.0)], "fine_census requires coarse_to_fine search"),
        (base + [("band", 16.0)], "band requires coarse_to_fine search"),
    ]
    for settings, reason in cases:
        reply = context.request(protocol.configure(7600, g, [("disparity", settings)]))
        if reason is None:
            assert reply is True
            context.request(protocol.release(7600))
        else:
            assert reply.reason == reason, settings
    odd = {"width": 321, "height": 96, "frame_width": 321, "frame_height": 96}
    reply = context.request(protocol.configure(7601, odd, [("disparity", base + [("search", 1.0)])]))
    assert reply.reason == "coarse_to_fine needs an even width and height, got 321x96"


def pack(image):
    """packing.ex:16-21: a u8 image's bytes as u32 words, four pixels a word."""
    image = np.ascontiguousarray(image, np.uint8)
    return image.view("<u4").reshape(image.shape[0], image.shape[1] // 4)


@pytest.mark.parametrize("frame_size,size", [((1456, 1088), (728, 544)), ((1000, 700), (728, 544)),
                                             ((333, 217), (100, 64)), ((320, 96), (320, 96))])
def test_decode_stage(context, frame_size, size):
    # stage/decode.ex: both frames area-resized to the working size (image.ex:402-476).
    (fw, fh), (w, h) = frame_size, size
    g = {"width": w, "height": h, "frame_width": fw, "frame_height": fh}
    frames = [scenes.noise(fw, fh, 3), scenes.noise(fw, fh, 4)]
    seg = build(context, [("decode", stages.Decode())], g)
    out = run_ok(seg, {"left_frame": frames[0], "right_frame": frames[1]})
    for side, frame in zip(("left", "right"), frames):
        same_bits(out[side], oracle.resize_area(frame, (h, w)))
    if fw % 4 == 0:
        # The same frames packed as u32 words (a pipeline without byte storage).
        packed = run_ok(seg, {"left_frame": pack(frames[0]), "right_frame": pack(frames[1])})
        for side in ("left", "right"):
            same_bits(packed[side], out[side])


def test_packed_images_enter_as_the_images_they_hold(context):
    left, right, _ = scenes.planar_pair(320, 96, seed=9, low=4, high=56)
    geometry, _ = rig_geometry(320, 96)
    seg = build(context, [("disparity", stages.Disparity(num_disparities=64, paths=8))], geometry)
    plain = run_ok(seg, {"left": left, "right": right})
    packed = run_ok(seg, {"left": pack(left), "right": pack(right)})
    same_bits(packed["disparity"], plain["disparity"])
    same_bits(plain["disparity"], oracle.sgm(left, right, 64, 8))


def test_decode_then_coarse_to_fine_from_packed_frames(context):
    # A byte-storage-free pipeline's shape: packed camera frames, Decode, then the matcher.
    w, h = 320, 96
    frames = [scenes.noise(2 * w, 2 * h, 5), scenes.noise(2 * w, 2 * h, 6)]
    g = {"width": w, "height": h, "frame_width": 2 * w, "frame_height": 2 * h}
    disparity = stages.Disparity(num_disparities=128, search="coarse_to_fine")
    seg = build(context, [("decode", stages.Decode()), ("disparity", disparity)], g)
    out = run_ok(seg, {"left_frame": pack(frames[0]), "right_frame": pack(frames[1])})
    left, right = (oracle.resize_area(f, (h, w)) for f in frames)
    same_bits(out["disparity"], oracle.coarse_to_fine(left, right, positive=True))


def test_decode_refuses_to_grow(context):
    g = {"width": 64, "height": 48, "frame_width": 64, "frame_height": 40}
    reply = context.request(protocol.configure(7700, g, [("decode", [])]))
    assert reply.reason == "resize_area only shrinks, got 40x64 to 48x64"
    reply = context.request(protocol.configure(7700, g, [("decode", [("kernel_size", 3.0)])]))
    assert reply.reason == "decode has no setting kernel_size"


def test_temporal_filter_carries_across_pairs(context):
    geometry, _ = rig_geometry(64, 32)
    seg = build(context, [("temporal_filter", stages.TemporalFilter(alpha=0.5, delta=2.0, hold=1))],
                geometry)
    want = oracle.Temporal((32, 64))
    rng = np.random.default_rng(5)
    base = rng.uniform(1, 50, (32, 64)).astype(np.float32)
    for pair in range(5):
        entry = np.where(rng.random((32, 64)) < 0.3, -1.0, base + rng.uniform(-3, 3, (32, 64))).astype(np.float32)
        out = run_ok(seg, {"disparity": entry})
        same_bits(out["disparity"], want.filter(entry, 0.5, 2.0, 1))


def test_stateful_pipeline_closed_forms(context):
    """stage/stateful_pipeline_test.exs:59-99 on a uniform image: the temporal filter's
    output over (10, 12, 12, -1, -1, -1, 50) and the quality jitter."""
    geometry, _ = rig_geometry(16, 8)
    members = [("temporal_filter", stages.TemporalFilter(alpha=0.5, delta=2.0, hold=2)),
               ("quality", stages.Quality())]
    seg = build(context, members, geometry)
    filtered, jitter = [], []
    for value in (10, 12, 12, -1, -1, -1, 50):
        out = run_ok(seg, {"disparity": np.full((8, 16), value, np.float32)})
        filtered.append(float(out["disparity"][0, 0]))
        sums = out["quality_sums"]
        jitter.append(float(sums[4] / sums[3]) if sums[3] else 0.0)
    assert filtered == [10.0, 11.0, 11.5, 11.5, 11.5, -1.0, 50.0]
    assert jitter[:3] == [0.0, 1.0, 0.5]
    seg.reload(seg.members, geometry)
    out = run_ok(seg, {"disparity": np.full((8, 16), 30.0, np.float32)})
    assert float(out["disparity"][0, 0]) == 30.0, "a reload restarts the carries"
    assert float(out["quality_sums"][3]) == 0.0


def test_quality_of_a_tall_image(context):
    """A 1200-row image: each column's totals gather 19 chunks of rows, added from blocks
    anywhere on the device. The sums are the oracle's (each column's sum is exact), and two
    segments fed the same frames give the same bits."""
    geometry, _ = rig_geometry(130, 1200)
    rng = np.random.default_rng(11)
    frames = [np.where(rng.random((1200, 130)) < 0.3, -1.0,
                       rng.uniform(0.5, 250.0, (1200, 130))).astype(np.float32) for _ in range(4)]
    quality = oracle.Quality((1200, 130))
    got = []
    for _ in range(2):
        seg = build(context, [("quality", stages.Quality())], geometry)
        got.append([run_ok(seg, {"disparity": f})["quality_sums"] for f in frames])
        seg.release()
    for g, f in zip(got[0], frames):
        np.testing.assert_allclose(g, quality.sums(f), rtol=1e-6)
    for a, b in zip(*got):
        same_bits(a, b)


def test_full_pipeline_728x544x128(context):
    """The benchmark's workload, every output against the oracle's chain: the rectified
    images, the disparity, depth_mm, the cloud grid and its kept count, and the quality sums
    of two pairs (the second against the first's carried map)."""
    geometry, cal = rig_geometry(728, 544)
    left_frame, right_frame, _ = scenes.raw_frames(cal, 7)
    members = [("rectification", stages.Rectification()),
               ("disparity", stages.Disparity(num_disparities=128, paths=8)),
               ("speckle_filter", stages.SpeckleFilter()), ("depth", stages.Depth()),
               ("point_cloud", stages.PointCloud(decimation=4)), ("quality", stages.Quality())]
    seg = build(context, members, geometry)
    out = run_ok(seg, {"left_frame": left_frame, "right_frame": right_frame})
    assert list(out) == ["left_frame", "right_frame", "left", "right", "disparity", "depth_mm",
                         "cloud_grid", "cloud_points", "quality_sums"]
    cams = seg.members[0][1].params(geometry)
    rect = {}
    for side, frame in (("left", left_frame), ("right", right_frame)):
        c = cams[side]
        rect[side] = oracle.rectify(frame, c["k"], c["d"], c["inverse_pr"], (544, 728))
        same_bits(out[side], rect[side])
    d = oracle.despeckle(oracle.sgm(rect["left"], rect["right"], 128, 8), 12, 1.0)
    same_bits(out["disparity"], d)
    assert (out["disparity"] > 0).mean() > 0.3, "the scene is mostly matched"
    f, b = geometry["focal_length_px"], geometry["baseline_m"]
    mm, m = oracle.depth(d, f, b)
    same_bits(out["depth_mm"], mm)
    cx, cy = geometry["principal_point"]
    points, kept = oracle.cloud(m, 4, f, cx, cy)
    assert out["cloud_grid"].shape == (136, 182, 3)
    same_bits(out["cloud_grid"], points)
    assert int(out["cloud_points"]) == kept and kept > 0
    # Quality against the previous pair's map: a second pair (the frames one pixel apart)
    # gives the both-matched and jitter terms something to count.
    want_first = oracle.Quality((544, 728))
    np.testing.assert_allclose(out["quality_sums"], want_first.sums(d), rtol=1e-5)
    moved = {"left_frame": np.roll(left_frame, 1, 1), "right_frame": np.roll(right_frame, 1, 1)}
    out2 = run_ok(seg, moved)
    rect2 = {side: oracle.rectify(moved[f"{side}_frame"], cams[side]["k"], cams[side]["d"],
                                  cams[side]["inverse_pr"], (544, 728)) for side in ("left", "right")}
    d2 = oracle.despeckle(oracle.sgm(rect2["left"], rect2["right"], 128, 8), 12, 1.0)
    same_bits(out2["disparity"], d2)
    sums = want_first.sums(d2)
    assert sums[3] > 0, "both maps match somewhere"
    np.testing.assert_allclose(out2["quality_sums"], sums, rtol=1e-5)


def test_refusals_map_like_the_port(context):
    geometry, _ = rig_geometry(32, 24)
    g = {k: geometry[k] for k in ("width", "height", "frame_width", "frame_height")}
    reply = context.request(protocol.configure(9001, g, [("colorize", [])]))
    assert reply.reason == "no HIP implementation of colorize"
    reply = context.request(protocol.run(9001, [], []))
    assert reply.reason == "segment 9001 is not configured"
    reply = context.request(protocol.configure(9001, g, [("temporal_filter", [("max_missed", 255.0),
                                                                              ("hold", 2.0)])]))
    assert reply.reason == "temporal_filter has no setting hold"
    reply = context.request(protocol.configure(9001, g, [("speckle_filter", [])]))
    assert reply.reason == "speckle_filter needs its radius setting"
    reply = context.request(protocol.configure(9001, {**g, "frame_height": 12}, [("rectification", [])]))
    assert reply.reason.startswith("rectification only shrinks: 32x12 frames")
    assert context.request(protocol.configure(9001, g, [("depth", [])])) is True
    reply = context.request(protocol.run(9001, [[]], [("disparity", np.zeros((2, 2), np.float32))]))
    assert "disparity is a [2, 2] tensor of type 2, expected a [24, 32] f32 map" in reply.reason
    reply = context.request(protocol.run(9001, [[]], []))
    assert reply.reason == "no disparity among the segment's tensors"  # the map is read first
    reply = context.request(protocol.run(9001, [[]], [("disparity", np.zeros((24, 32), np.float32))]))
    assert reply.reason == "no focal_length_px among the stage's inputs"
    assert isinstance(protocol.decode(bytes([9])), protocol.Refused)
    reply = context.request(bytes([2, 1, 0]))
    assert "ends 2 bytes early" in reply.reason
    reply = context.request(bytes([9, 1, 0, 0, 0]))
    assert reply.reason == "unknown request kind 9"
    assert context.request(protocol.release(9001)) is None
    assert context.request(protocol.run(9001, [[]], [])).reason == "segment 9001 is not configured"


def test_a_refused_reconfigure_leaves_the_segment_unconfigured(context):
    g = {"width": 32, "height": 24, "frame_width": 32, "frame_height": 24}
    assert context.request(protocol.configure(77, g, [("depth", [])])) is True
    assert isinstance(context.request(protocol.configure(77, g, [("colorize", [])])), protocol.Refused)
    assert context.request(protocol.run(77, [[]], [])).reason == "segment 77 is not configured"


def test_penalties_must_be_integer_bytes(context):
    geometry, _ = rig_geometry(32, 24)
    left = scenes.noise(32, 24, 1)
    seg = build(context, [("disparity", stages.Disparity())], geometry)
    seg.inputs = [[("p1", np.array([8.5], np.float32)), ("p2", np.array([64.0], np.float32)),
                   ("uniqueness", np.array([1.0], np.float32))]]
    result = seg.await_(seg.send_request({"left": left, "right": left}))
    assert result == ("error", "p1 must be an integer from 0 to 255, got 8.5")


def test_pairs_complete_in_the_order_sent(context):
    """native_segment.ex:167-171: carried state depends on the order."""
    geometry, _ = rig_geometry(48, 16)
    seg = build(context, [("temporal_filter", stages.TemporalFilter(alpha=0.5, delta=8.0, hold=0))],
                geometry)
    values = [8.0, 12.0, 6.0, 11.0, 14.0]
    requests = [seg.send_request({"disparity": np.full((16, 48), v, np.float32)}) for v in values]
    got = [float(seg.await_(r)[1]["disparity"][0, 0]) for r in requests]
    want, previous = [], None
    for v in values:
        previous = v if previous is None else 0.5 * v + 0.5 * previous
        want.append(previous)
    assert got == want


def test_device_entries_run_in_place(context):
    """nxs_run_device: entries already in HBM, outputs left in the segment's memory."""
    from nx_stereo_amd import hip

    geometry, _ = rig_geometry(96, 64)
    left, right = scenes.scene(96, 64, 11)
    seg = build(context, [("disparity", stages.Disparity())], geometry)
    assert context.request(seg.configure_record()) is True
    seg.configured_version = seg.version
    dl, dr = hip.Buffer.of(left), hip.Buffer.of(right)
    outs = context.run_device(seg.run_record({}), [("left", 0, (64, 96), dl.ptr),
                                                   ("right", 0, (64, 96), dr.ptr)])
    assert [(o[0], o[1], o[2]) for o in outs] == [("disparity", 2, (64, 96))]
    hip.synchronize()
    same_bits(hip.copy_to_host(outs[0][3], (64, 96), np.float32), oracle.sgm(left, right, 64, 8))


TAIL_CHAINS = [
    ["speckle_filter", "temporal_filter", "depth", "point_cloud", "quality"],
    ["speckle_filter", "depth", "point_cloud", "quality"],
    ["temporal_filter", "quality"],
    ["depth", "point_cloud"],
    ["speckle_filter", "quality"],
    ["quality", "depth", "point_cloud"],  # quality first: it runs alone, the rest fused
    ["depth", "quality", "point_cloud"],  # the cloud after quality: not in the fused run
]


def tail_member(name, decimation):
    return {"speckle_filter": lambda: stages.SpeckleFilter(min_support=9, max_diff=1.5),
            "temporal_filter": lambda: stages.TemporalFilter(alpha=0.4, delta=2.0, hold=2),
            "depth": lambda: stages.Depth(),
            "point_cloud": lambda: stages.PointCloud(decimation=decimation),
            "quality": lambda: stages.Quality()}[name]()


@pytest.mark.parametrize("chain", TAIL_CHAINS, ids=["-".join(c) for c in TAIL_CHAINS])
@pytest.mark.parametrize("size,decimation", [((200, 77), 3), ((130, 49), 8), ((64, 16), 4)])
def test_fused_tail_equals_the_stages_one_by_one(context, chain, size, decimation, monkeypatch):
    """Consecutive per-pixel stages run as one fused launch (and a finishing block); with
    NXS_FUSE_TAIL=0 at configure each runs alone. Both give the same bits on three pairs in a
    row (the carries), at sizes off the 64x16 tiles and decimations that put two grid rows in
    one thread (8) or none (3), and both equal the oracle's chain."""
    w, h = size
    geometry, _ = rig_geometry(w, h)
    rng = np.random.default_rng(w + h + decimation)
    maps = []
    for k in range(3):
        m = np.where(rng.random((h, w)) < 0.25, -1.0, rng.uniform(0.3, 60.0, (h, w))).astype(np.float32)
        m[h // 4:h // 2, w // 4:w // 2] = 9.5 + 0.25 * k
        maps.append(m)
    outs = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("NXS_FUSE_TAIL", fuse)
        seg = build(context, [(name, tail_member(name, decimation)) for name in chain], geometry)
        outs[fuse] = [run_ok(seg, {"disparity": m}) for m in maps]
        seg.release()
    f, b = geometry["focal_length_px"], geometry["baseline_m"]
    cx, cy = geometry["principal_point"]
    temporal, quality = oracle.Temporal((h, w)), oracle.Quality((h, w))
    for k, m in enumerate(maps):
        fused, alone = outs["1"][k], outs["0"][k]
        assert list(fused) == list(alone)
        for name in fused:
            same_bits(fused[name], alone[name])
        d = m.copy()
        mm = metres = None
        for name in chain:
            if name == "speckle_filter":
                d = oracle.despeckle(d, 9, 1.5)
            elif name == "temporal_filter":
                d = temporal.filter(d, 0.4, 2.0, 2)
            elif name == "depth":
                mm, metres = oracle.depth(d, f, b)
            elif name == "point_cloud":
                points, kept = oracle.cloud(metres, decimation, f, cx, cy)
                same_bits(fused["cloud_grid"], points)
                assert int(fused["cloud_points"]) == kept
            elif name == "quality":
                np.testing.assert_allclose(fused["quality_sums"], quality.sums(d), rtol=1e-5)
        if "disparity" in fused:
            same_bits(fused["disparity"], d)
        if mm is not None:
            same_bits(fused["depth_mm"], mm)
