"""Serving many camera streams through one context.

Each stream is a HipSegment at its own size (Disparity -> TemporalFilter ->
Quality by default), configured once; its rectified pairs arrive by a seeded
Poisson process at the stream's rate, from a planar scene drifting half a pixel
per frame. A submitting thread sends each pair's run record at its arrival
time without waiting (nxs_submit with a completion callback). The context's
worker admits a pair as soon as its segment has a free run slot and the
matcher workspace pool's HBM budget allows; its work then overlaps the pairs
of other streams on the device, and the completion thread calls back when its
last download lands. Latency is completion minus arrival, as
test/bench/stream.exs:208-225 measures it per frame.

`serve` is that continuous form. `serve_one_at_a_time` is the plain form for
comparison: the same arrivals, one pair in flight at a time in arrival
order, as a Port-driven NativeSegment computes them.
"""
from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native, protocol, scenes, stages
from .segment import Context, HipSegment


@dataclass
class StreamSpec:
    """One camera stream: its rectified size, disparity range, frame rate and seed."""

    width: int
    height: int
    disparities: int
    rate_hz: float
    seed: int
    members: list = field(default_factory=list)  # (name, stage); default: disparity, temporal, quality

    def stage_list(self):
        if self.members:
            return self.members
        return [("disparity", stages.Disparity(num_disparities=self.disparities, paths=8)),
                ("temporal_filter", stages.TemporalFilter(alpha=0.4, delta=2.0, hold=2)),
                ("quality", stages.Quality())]


def stream_frames(spec: StreamSpec, count: int):
    """`count` rectified (left, right) pairs of a planar scene drifting 0.5 px per frame.

    One planar pair wider than the stream by the drift is generated; frame k views it
    from column floor(k / 2), so the scene moves one pixel every two frames.
    """
    drift = (count + 1) // 2
    left, right, _ = scenes.planar_pair(spec.width + drift, spec.height, seed=spec.seed, low=4,
                                        high=max(spec.disparities - 8, 8))
    frames = []
    for k in range(count):
        s = k // 2
        frames.append((np.ascontiguousarray(left[:, s:s + spec.width]),
                       np.ascontiguousarray(right[:, s:s + spec.width])))
    return frames


def poisson_arrivals(rate_hz: float, duration_s: float, seed: int) -> np.ndarray:
    """Arrival times in [0, duration) of a Poisson process at `rate_hz`."""
    rng = np.random.default_rng(seed)
    times, t = [], rng.exponential(1.0 / rate_hz)
    while t < duration_s:
        times.append(t)
        t += rng.exponential(1.0 / rate_hz)
    return np.array(times)


def geometry_of(spec: StreamSpec) -> dict:
    return {"width": spec.width, "height": spec.height, "frame_width": spec.width,
            "frame_height": spec.height, "focal_length_px": 873.0, "baseline_m": 0.2,
            "principal_point": (spec.width / 2.0, spec.height / 2.0)}


def configure_streams(context: Context, specs) -> list[HipSegment]:
    """One configured segment per stream."""
    segments = []
    for spec in specs:
        ok, seg = HipSegment.build(spec.stage_list(), geometry_of(spec), context)
        if ok != "ok":
            raise RuntimeError(f"stream refused: {seg}")
        reply = context.request(seg.configure_record())
        if reply is not True:
            raise RuntimeError(f"configure refused: {reply}")
        seg.configured_version = seg.version
        segments.append(seg)
    return segments


@dataclass
class Served:
    """Per pair, in submission order: stream, frame index, arrival and completion
    (seconds from the trace's start), and the reply when kept."""

    stream: np.ndarray
    frame: np.ndarray
    arrival: np.ndarray
    done: np.ndarray
    replies: list

    @property
    def latency_ms(self) -> np.ndarray:
        return (self.done - self.arrival) * 1e3

    def summary(self) -> dict:
        """pairs/s over the trace, and latency percentiles by nearest rank as
        test/bench/stream.exs:208-225 reports them (the value at ceil(n p) - 1 of the sorted
        latencies), with the count over its 33.3 ms frame budget."""
        lat = self.latency_ms
        span = float(self.done.max() - self.arrival.min())
        return {"pairs": int(len(lat)), "pairs_per_s": round(len(lat) / span, 2),
                "p50_ms": round(nearest_rank(lat, 0.5), 3),
                "p90_ms": round(nearest_rank(lat, 0.9), 3),
                "p99_ms": round(nearest_rank(lat, 0.99), 3),
                "max_ms": round(float(lat.max()), 3),
                "over_budget": int((lat > BUDGET_MS).sum())}


# stream.exs:37: a frame's budget at 30 Hz.
BUDGET_MS = 1000.0 / 30


def nearest_rank(values, fraction: float) -> float:
    """stream.exs:223: the sorted values' element ceil(n * fraction) - 1."""
    ordered = np.sort(np.asarray(values, np.float64))
    return float(ordered[max(int(np.ceil(len(ordered) * fraction)) - 1, 0)])


def _schedule(arrivals):
    """(time, stream, frame) over every stream's arrivals, by time."""
    order = [(t, s, k) for s, ts in enumerate(arrivals) for k, t in enumerate(ts)]
    order.sort()
    return order


def serve(context: Context, segments, records, arrivals, keep_replies: bool = False,
          all_at_once: bool = False) -> Served:
    """Sends records[s][k] at arrivals[s][k] seconds (all at once with `all_at_once`)
    and collects every completion through the context's callback."""
    order = _schedule(arrivals)
    n = len(order)
    done = np.zeros(n)
    replies = [None] * n
    tags = np.zeros(n, np.int32)
    remaining = threading.Semaphore(0)
    start = [0.0]

    @_native.DONE_FN
    def on_done(user, ticket, reply, length):
        i = int(user or 0)
        done[i] = time.perf_counter() - start[0]
        tags[i] = reply[0] if length else 0
        if keep_replies:
            replies[i] = ctypes.string_at(reply, length)
        remaining.release()

    # Records as ctypes buffers up front: submitting copies them, nothing else. A stream
    # with fewer records than arrivals sends them round again.
    made = {}
    buffers = []
    for _, s, k in order:
        key = (s, k % len(records[s]))
        if key not in made:
            rec = records[s][key[1]]
            made[key] = (ctypes.c_uint8 * len(rec)).from_buffer_copy(rec)
        buffers.append(made[key])
    lib = context.lib
    start[0] = time.perf_counter()
    for i, (t, s, k) in enumerate(order):
        if not all_at_once:
            wait = t - (time.perf_counter() - start[0])
            if wait > 0:
                time.sleep(wait)
        ticket = lib.nxs_submit(context.handle, buffers[i], len(buffers[i]), on_done,
                                ctypes.c_void_p(i))
        if ticket < 0:
            raise _native.NativeError(f"nxs_submit failed: {_native.last_error()}")
    for _ in range(n):
        if not remaining.acquire(timeout=120):
            raise TimeoutError("a served pair did not complete")
    if (tags != protocol.TAG_OUTPUTS).any():
        bad = int(np.argmax(tags != protocol.TAG_OUTPUTS))
        raise RuntimeError(f"pair {bad} was refused: {replies[bad] or tags[bad]}")
    arrival = np.zeros(n) if all_at_once else np.array([t for t, _, _ in order])
    return Served(np.array([s for _, s, _ in order]), np.array([k for _, _, k in order]),
                  arrival, done, replies)


def serve_one_at_a_time(context: Context, segments, records, arrivals,
                        all_at_once: bool = False) -> Served:
    """The plain form: pairs in arrival order, each waited for before the next is sent."""
    order = _schedule(arrivals)
    n = len(order)
    done = np.zeros(n)
    start = time.perf_counter()
    for i, (t, s, k) in enumerate(order):
        if not all_at_once:
            wait = t - (time.perf_counter() - start)
            if wait > 0:
                time.sleep(wait)
        ticket = context.submit(records[s][k % len(records[s])])
        context.wait(ticket)
        reply = context.reply(ticket)
        done[i] = time.perf_counter() - start
        if not reply or reply[0] != protocol.TAG_OUTPUTS:
            raise RuntimeError(f"pair {i} was refused: {reply!r}")
    arrival = np.zeros(n) if all_at_once else np.array([t for t, _, _ in order])
    return Served(np.array([s for _, s, _ in order]), np.array([k for _, _, k in order]),
                  arrival, done, [None] * n)


def run_records(segments, frames):
    """records[s][k]: stream s's k-th pair as a run record."""
    return [[seg.run_record({"left": left, "right": right}) for left, right in fs]
            for seg, fs in zip(segments, frames)]


def load_sweep(context: Context, segments, records, specs, factors, duration_s: float,
               seed: int = 0) -> list[dict]:
    """Offered load against latency, both ways: for each factor, every stream's Poisson
    arrivals at factor x its rate for `duration_s`, served continuously and one pair at a
    time (the Port's way); pairs/s carried and nearest-rank p50/p90/p99 for each."""
    points = []
    for factor in factors:
        arrivals = [poisson_arrivals(spec.rate_hz * factor, duration_s, spec.seed + seed + 7919 * i)
                    for i, spec in enumerate(specs)]
        offered = sum(len(a) for a in arrivals) / duration_s
        point = {"factor": factor, "offered_pairs_per_s": round(offered, 1)}
        point["continuous"] = serve(context, segments, records, arrivals).summary()
        point["one_at_a_time"] = serve_one_at_a_time(context, segments, records, arrivals).summary()
        points.append(point)
    return points


def at_equal_p99(points, target_ms: float) -> dict:
    """The most pairs/s each form carried in the sweep with its p99 latency within
    `target_ms`, and the offered load it carried it at."""
    out = {"p99_target_ms": round(target_ms, 3)}
    for form in ("continuous", "one_at_a_time"):
        ok = [p for p in points if p[form]["p99_ms"] <= target_ms]
        best = max(ok, key=lambda p: p[form]["pairs_per_s"]) if ok else None
        out[form] = ({"pairs_per_s": best[form]["pairs_per_s"], "p99_ms": best[form]["p99_ms"],
                      "offered_pairs_per_s": best["offered_pairs_per_s"]} if best else None)
    return out
