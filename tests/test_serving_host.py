"""Host side of the served form, without a GPU: arrivals, the merged schedule, drifting frames."""
import numpy as np

from nx_stereo_amd import serving


def test_poisson_arrivals_rate_and_order():
    t = serving.poisson_arrivals(60.0, 50.0, seed=3)
    assert (np.diff(t) > 0).all() and t[0] >= 0 and t[-1] < 50.0
    assert abs(len(t) / 50.0 - 60.0) < 4 * np.sqrt(60.0 / 50.0) * 3, len(t)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 60.0) < 0.1 / 60.0
    # Exponential gaps: the coefficient of variation is 1.
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1
    assert np.array_equal(t, serving.poisson_arrivals(60.0, 50.0, seed=3)), "seeded"


def test_schedule_merges_streams_by_time_keeping_each_stream_in_order():
    arrivals = [np.array([0.1, 0.5, 0.9]), np.array([0.2, 0.3]), np.array([0.05])]
    order = serving._schedule(arrivals)
    assert [s for _, s, _ in order] == [2, 0, 1, 1, 0, 0]
    for s in range(3):
        assert [k for _, t, k in order if t == s] == list(range(len(arrivals[s])))


def test_frames_drift_half_a_pixel_per_frame():
    spec = serving.StreamSpec(40, 12, 64, 30.0, seed=9)
    frames = serving.stream_frames(spec, 6)
    assert all(l.shape == (12, 40) and r.shape == (12, 40) for l, r in frames)
    for k in range(0, 4, 2):
        (l0, r0), (l1, r1), (l2, r2) = frames[k], frames[k + 1], frames[k + 2]
        assert np.array_equal(l0, l1) and np.array_equal(r0, r1)
        # Two frames on, the view has moved one column.
        assert np.array_equal(l2[:, :-1], l0[:, 1:]) and np.array_equal(r2[:, :-1], r0[:, 1:])


def test_served_summary():
    s = serving.Served(np.zeros(4, int), np.arange(4), np.array([0.0, 0.1, 0.2, 0.3]),
                       np.array([0.01, 0.12, 0.23, 0.34]), [None] * 4)
    out = s.summary()
    assert out["pairs"] == 4
    # Nearest rank (stream.exs:223): of 10, 20, 30, 40 ms, p50 is element ceil(2) - 1 = 20 ms
    # and p90 and p99 element ceil(3.6) - 1 = ceil(3.96) - 1 = 40 ms.
    assert abs(out["p50_ms"] - 20.0) < 1e-6 and abs(out["p90_ms"] - 40.0) < 1e-6
    assert abs(out["p99_ms"] - 40.0) < 1e-6 and out["max_ms"] == 40.0
    assert out["over_budget"] == 1  # only 40 ms is over the 33.3 ms budget
    assert abs(out["pairs_per_s"] - 4 / 0.34) < 0.01


def test_nearest_rank_percentiles():
    values = np.arange(1, 101, dtype=float)[::-1]  # 100 .. 1, unsorted
    assert serving.nearest_rank(values, 0.5) == 50.0
    assert serving.nearest_rank(values, 0.9) == 90.0
    assert serving.nearest_rank(values, 0.99) == 99.0
    assert serving.nearest_rank([7.0], 0.99) == 7.0
    assert serving.nearest_rank([3.0, 1.0, 2.0], 0.5) == 2.0  # ceil(1.5) - 1 = 1


def test_at_equal_p99_picks_the_most_carried_within_the_target():
    def point(offered, c, cp99, o, op99):
        return {"offered_pairs_per_s": offered,
                "continuous": {"pairs_per_s": c, "p99_ms": cp99},
                "one_at_a_time": {"pairs_per_s": o, "p99_ms": op99}}
    sweep = [point(400, 400, 1.0, 400, 2.0), point(800, 800, 2.0, 780, 6.0),
             point(1600, 1590, 4.0, 1200, 300.0), point(2400, 2100, 90.0, 1300, 900.0)]
    out = serving.at_equal_p99(sweep, 5.0)
    assert out["continuous"] == {"pairs_per_s": 1590, "p99_ms": 4.0, "offered_pairs_per_s": 1600}
    assert out["one_at_a_time"] == {"pairs_per_s": 400, "p99_ms": 2.0, "offered_pairs_per_s": 400}
    assert serving.at_equal_p99(sweep, 0.5)["continuous"] is None
