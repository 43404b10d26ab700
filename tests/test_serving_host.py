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
    assert abs(out["p50_ms"] - 25.0) < 1e-6 and out["max_ms"] == 40.0
    assert abs(out["pairs_per_s"] - 4 / 0.34) < 0.01
