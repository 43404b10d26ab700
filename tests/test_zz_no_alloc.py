"""nxs_match_device_noalloc: the matcher with the library's static scratch as its workspace.
Its output equals nxs_match_device's (with a workspace of the caller's) bit for bit, on two
calls in a row on one stream, the second smaller than the first (the scratch keeps the first
call's bytes), and both equal the CPU checker."""
import numpy as np
import pytest

from nx_stereo_amd import _native, hip, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

UNIQUENESS = 1.0526315789473684


def test_no_alloc_matcher_equals_the_allocating_one_on_two_calls():
    cases = [(480, 270, 64, 8, 8, 64, 31), (364, 272, 64, 3, 10, 70, 32)]
    stream = hip.Stream()
    try:
        for w, h, d, paths, p1, p2, seed in cases:
            assert _native.workspace_size(w, h, d, paths) <= _native.NOALLOC_SCRATCH_BYTES
            left, right, _ = scenes.planar_pair(w, h, seed=seed, low=4, high=d - 8)
            print(f"no-alloc matcher {w}x{h}x{d}, {paths} paths", flush=True)
            l, r = hip.Buffer.of(left), hip.Buffer.of(right)
            got, want = hip.Buffer(4 * w * h), hip.Buffer(4 * w * h)
            ws = hip.Buffer(_native.workspace_size(w, h, d, paths))
            _native.match_device_noalloc(l.ptr, r.ptr, w, h, d, paths, p1, p2, UNIQUENESS,
                                         got.ptr, stream.handle)
            _native.match_device(l.ptr, r.ptr, w, h, d, paths, p1, p2, UNIQUENESS, want.ptr,
                                 ws.ptr, ws.nbytes, stream.handle)
            stream.synchronize()
            a = got.download((h, w), np.float32)
            b = want.download((h, w), np.float32)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (w, h, d, paths)
            c = oracle.cpu_sgm(left, right, d, paths, p1, p2, UNIQUENESS)
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), (w, h, d, paths)
            assert (a > 0).mean() > 0.3
            for buf in (l, r, got, want, ws):
                buf.free()
    finally:
        stream.destroy()


def test_sizes_beyond_the_scratch_are_refused():
    w, h, d = 728, 544, 128
    assert _native.workspace_size(w, h, d, 8) > _native.NOALLOC_SCRATCH_BYTES
    l, o = hip.Buffer(w * h), hip.Buffer(4 * w * h)
    try:
        with pytest.raises(_native.NativeError, match="use nxs_match_device"):
            _native.match_device_noalloc(l.ptr, l.ptr, w, h, d, 8, 8, 64, UNIQUENESS, o.ptr)
    finally:
        l.free()
        o.free()
