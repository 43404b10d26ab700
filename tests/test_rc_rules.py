"""A CPU model of how path_rc_kernel's two waves meet by recomputation (sgm.hip).

Both waves of a block walk the same lines, padded to T = S * ceil(longest / S) positions:
the forward wave steps through positions 0, 1, ..., the backward wave through T-1, T-2, ...
A line of length L has its pixels at positions 0 .. L-1; beyond them costs load as zeros
and stores are dropped. A wave's phase 1 is its first h1 steps (ceil(n/2) segments of S for
the forward wave, floor(n/2) for the backward one), where it walks its own path and, at the
end of every segment but the last, stores the path's bytes at that step's pixel (a
checkpoint). After one barrier, phase 2 takes the rest of the line S steps at a time: the
wave restores the checkpoint at the pixel just past the segment (or the entering path, zero
bytes, for the line's last segment), recomputes the other direction's path backward over
the segment's costs, then walks its own steps and stores the byte sums.

The model replays that schedule with the kernel's own index formulas, one line of a block
at a time, and checks:

* every load of a checkpoint finds one the other wave stored in its phase 1 and nobody
  has overwritten since, and every store lands on a real pixel of the line;
* each pixel's pair sum is stored exactly once;
* the stored pair sums equal the two directions' path costs of the whole line, computed
  directly (the oracle's recurrence, normalised, with the p2 sentinel at the band's ends).

Breaking the kernel's rules in the model (a checkpoint one step late, a segment restored
from the checkpoint of the segment after it, the other direction recomputed in the wrong
order) must fail it.
"""
import numpy as np
import pytest

P1, P2 = 8, 64


def step(prev, cost):
    """One step of a direction's path: L = C + min(prev, prev[d+-1] + p1, m + p2) - m."""
    m = prev.min()
    near = np.full_like(prev, 1 << 20)
    near[1:] = np.minimum(near[1:], prev[:-1])
    near[:-1] = np.minimum(near[:-1], prev[1:])
    return cost + np.minimum(np.minimum(prev, near + P1), m + P2) - m


def path(costs):
    """A direction's path costs over a whole line, entering with zeros."""
    out, state = [], np.zeros(costs.shape[1], np.int64)
    for c in costs:
        state = step(state, c)
        out.append(state)
    return np.array(out)


def replay(length, longest, S, costs, late_checkpoint=0, skip_segment=False, wrong_order=False):
    """The kernel's schedule for one line of a block whose longest line is `longest`;
    returns the pair sums it stores, by position."""
    T = S * -(-longest // S)
    n = T // S
    D = costs.shape[1]
    volume = {}       # position -> ("checkpoint", wave, state) or ("sum", values)
    stored_sums = {}  # position -> values
    waves = {}
    for wave in (0, 1):
        h1 = (-(-n // 2) if wave == 0 else n // 2) * S
        lo = 0 if wave == 0 else T - length

        def position(t, wave=wave):
            return t if wave == 0 else T - 1 - t

        def real(t, lo=lo):
            return 0 <= t - lo < length

        def cost(t, wave=wave, real=real, position=position):
            return costs[position(t)] if real(t) else np.zeros(D, np.int64)

        waves[wave] = (h1, position, real, cost)

    # Phase 1 of both waves (independent), checkpoints into the volume.
    states = {}
    for wave in (0, 1):
        h1, position, real, cost = waves[wave]
        state = np.zeros(D, np.int64)
        for t in range(h1):
            state = step(state, cost(t))
            end = t + 1 - late_checkpoint  # a checkpoint ends a segment of S steps
            if end % S == 0 and 0 < end < h1:
                if real(t):
                    assert position(t) not in volume
                    volume[position(t)] = ("checkpoint", wave, state.copy())
        states[wave] = state
    # Barrier; phase 2, each wave in its own program order.
    for wave in (0, 1):
        h1, position, real, cost = waves[wave]
        other = 1 - wave
        h2 = T - h1
        own = states[wave]
        for k in range(h2 // S):
            t0 = h1 + k * S
            last = (k + 1) * S >= h2
            if last:
                entering = np.zeros(D, np.int64)
            else:
                t = t0 + (2 * S if skip_segment else S)
                if real(t):
                    kind = volume.get(position(t))
                    assert kind is not None and kind[0] == "checkpoint" and kind[1] == other, \
                        (length, longest, S, wave, k)
                    entering = kind[2]
                else:
                    entering = np.zeros(D, np.int64)  # a masked load reads zeros
            rec, theirs = entering.copy(), {}
            for r in (range(S) if wrong_order else range(S - 1, -1, -1)):
                rec = step(rec, cost(t0 + r))
                theirs[r] = rec
            for p in range(S):
                own = step(own, cost(t0 + p))
                if real(t0 + p):
                    pos = position(t0 + p)
                    assert pos not in stored_sums, (length, longest, S, pos)
                    stored_sums[pos] = own + theirs[p]
                    volume[pos] = ("sum",)
    return stored_sums


def check(length, longest, S, seed, **faults):
    rng = np.random.default_rng(seed)
    D = 8
    costs = rng.integers(0, 49, size=(length, D)).astype(np.int64)
    sums = replay(length, longest, S, costs, **faults)
    assert sorted(sums) == list(range(length)), (length, longest, S)
    want = path(costs) + path(costs[::-1])[::-1]
    for pos in range(length):
        assert np.array_equal(sums[pos], want[pos]), (length, longest, S, pos)


@pytest.mark.parametrize("S", [2, 4, 8])
def test_every_line_length_up_to_70(S):
    for length in range(1, 71):
        check(length, length, S, seed=length)


@pytest.mark.parametrize("S", [2, 4, 8])
def test_lines_shorter_than_the_block_longest(S):
    """A diagonal block's lines differ in length: each is padded to the block's T, and the
    backward wave walks zero costs until it reaches the line's last pixel."""
    for longest in (1, 5, 8, 17, 33, 64):
        for length in {1, max(longest - 1, 1), max(longest - 3, 1), max(longest // 2, 1),
                       longest}:
            check(length, longest, S, seed=1000 * longest + length)


def test_zero_costs_keep_the_entering_path():
    state = np.zeros(16, np.int64)
    for _ in range(5):
        state = step(state, np.zeros(16, np.int64))
    assert not state.any()


def test_the_model_catches_a_late_checkpoint():
    with pytest.raises(AssertionError):
        for length in range(9, 40):
            check(length, length, 4, seed=length, late_checkpoint=1)


def test_the_model_catches_the_next_segment_s_checkpoint():
    with pytest.raises(AssertionError):
        for length in range(9, 40):
            check(length, length, 4, seed=length, skip_segment=True)


def test_the_model_catches_a_recomputation_in_the_wrong_order():
    with pytest.raises(AssertionError):
        for length in range(9, 40):
            check(length, length, 4, seed=length, wrong_order=True)
