"""A CPU model of how path2_kernel's two waves meet (sgm.hip, the two-wave aggregation).

One wave walks a block's four lines forward, the other backward. Each stores its own bytes
before a line's middle and, past it, adds the other's bytes of the same pixel: from an LDS
ring near the middle, from memory (loaded PREFETCH steps ahead) further on. The model
replays the kernel's schedule step by step, with the same constants and phase bounds, and
checks every read:

* a ring read needs the slot written by the other wave, a barrier since, and a slot that
  has not been rewritten;
* a memory read, issued PREFETCH steps early, needs a barrier at or before its issue at
  which the other's store was complete. At a barrier the wave has waited with
  ``s_waitcnt vmcnt(3 LAG - 1)``: every access but the last 3 LAG - 1 it issued is complete
  (accesses complete in issue order). Phase A steps issue a store and one load; the other
  steps a store and two loads.

A schedule with a wait that is too loose must fail the model, so the model can fail.
"""
import bisect

import pytest


def constants(prefetch):
    lag = 5 if prefetch == 4 else 6
    x = 2 * prefetch + lag + 2
    ring = 16 if x + 1 <= 16 else 32
    return lag, x, ring


def phases(lengths, prefetch, x):
    """The kernel's block-uniform bounds: phase A is [0, a_end), phase C [c_start, c_end)."""
    window_lo = min(lengths) - 1 - x
    window_hi = max(lengths) - 1 + x
    a_end = (max(0, (window_lo + 1) >> 1)) & ~(prefetch - 1)
    c_start = ((window_hi >> 1) + prefetch) & ~(prefetch - 1)
    c_end = min(lengths) & ~(prefetch - 1)
    return window_lo, window_hi, a_end, c_start, c_end


def check_block(lengths, prefetch, wait_slack=0):
    """Replays one block; raises AssertionError at the first read the rules do not cover."""
    lag, x, ring = constants(prefetch)
    window_lo, window_hi, a_end, c_start, c_end = phases(lengths, prefetch, x)
    steps = -(-max(lengths) // prefetch) * prefetch
    accesses = []  # (step, kind) in issue order
    barriers, complete_upto, ring_written = [], {}, set()

    def barrier(i):
        done = len(accesses) - (3 * lag - 1 + wait_slack)
        complete_upto[i] = max([s for s, kind in accesses[:max(done, 0)] if kind == "store"],
                               default=-1)
        barriers.append(i)

    for i in range(steps):
        in_window = window_lo <= 2 * i <= window_hi
        if i < a_end:
            for length in lengths:  # phase A needs no mask and reads nothing of the other's
                assert i < length and 2 * i - (length - 1) < 0
                assert 2 * (i + prefetch) - (length - 1) < x
            accesses += [(i, "store"), (i, "load")]
            continue
        if c_start <= i < c_end:
            for length in lengths:  # phase C needs no mask: the other's bytes are in memory
                assert i < length and 2 * i - (length - 1) >= x
            if i % prefetch == 0:
                barrier(i)
        else:
            if in_window:
                ring_written.add(i)
            if in_window or i % prefetch == 0:
                barrier(i)
        accesses += [(i, "store"), (i, "load"), (i, "load")]

    def last_barrier(t):
        k = bisect.bisect_right(barriers, t) - 1
        return barriers[k] if k >= 0 else None

    for length in lengths:
        for i in range(length):
            past = 2 * i - (length - 1)
            if past < 0:
                continue
            made = i - past  # the other wave's step for this pixel
            if past < x:
                assert made in ring_written, ("ring slot not written", lengths, i)
                b = last_barrier(i)
                assert b is not None and b >= made, ("no barrier after the ring write", lengths, i)
                assert past < ring - 1, ("ring slot rewritten", lengths, i)
            else:
                b = last_barrier(i - prefetch)
                assert b is not None and made <= complete_upto[b], \
                    ("store not complete", lengths, i, made, b)


# Line lengths in a block, relative to its first line's. Diagonal blocks hold unequal lines;
# the last block of an orientation may hold lines of length 0 (past the last line).
SHAPES = [(0, 0, 0, 0), (0, 1, 2, 3), (0, -1, -2, -3), (0, 1, 1, 2), (0, 5, 9, 3),
          (0, 0, 0, -1), (0, -300, -300, -300)]


def blocks(last):
    for first in range(1, last):
        for shape in SHAPES:
            yield [max(0, first + d) for d in shape]


@pytest.mark.parametrize("prefetch", [2, 4, 8])
def test_every_read_is_covered(prefetch):
    for lengths in blocks(260):
        check_block(lengths, prefetch)


def test_a_looser_wait_fails():
    """A wait that leaves five more steps' accesses outstanding leaves a read uncovered."""
    with pytest.raises(AssertionError):
        for prefetch in (2, 4, 8):
            for lengths in blocks(200):
                check_block(lengths, prefetch, wait_slack=15)
