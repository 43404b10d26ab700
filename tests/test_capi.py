"""The C ABI library loads and exports every entry point include/nxs_hip.h declares,
and the checks it makes before touching a device (no compute runs here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from nx_stereo_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nxs_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(nxs_[a-z_]+)\s*\(", text))
    return sorted(n for n in names if n != "nxs_done_fn")


def test_header_declares_the_entry_points_the_bindings_use():
    assert set(declared_functions()) == set(_native.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    for symbol in declared_functions():
        assert hasattr(lib, symbol), symbol


def test_workspace_size_grows_with_the_volume():
    small = _native.workspace_size(64, 48, 64, 8)
    assert small >= 64 * 48 * 64 * 5
    assert _native.workspace_size(64, 48, 128, 8) > small
    assert _native.workspace_size(64, 48, 64, 3) < small


@pytest.mark.parametrize("args,reason", [
    ((8, 48, 64, 8, 8, 64, 1.0), "wider than 8"),
    ((64, 6, 64, 8, 8, 64, 1.0), "higher than 6"),
    ((4097, 48, 64, 8, 8, 64, 1.0), "at most 4096"),
    ((64, 48, 96, 8, 8, 64, 1.0), "disparities must be"),
    ((64, 48, 64, 4, 8, 64, 1.0), "paths must be 3 or 8"),
    ((64, 48, 64, 8, 64, 64, 1.0), "p1 < p2"),
    ((64, 48, 64, 8, -1, 64, 1.0), "0 <= p1"),
    ((64, 48, 64, 8, 8, 80, 1.0), "p1 < p2"),
    ((64, 48, 64, 8, 8, 64, 0.9), "uniqueness"),
    ((64, 48, 64, 8, 8, 64, float("nan")), "uniqueness"),
])
def test_match_refuses_what_the_reference_refuses_before_any_device_work(args, reason):
    """lib.rs:41-73 and 645-687: Parameters::check."""
    w, h, d, paths, p1, p2, u = args
    left = np.zeros((h, w), np.uint8)
    with pytest.raises(_native.NativeError, match=reason):
        _native.match_host(left, left, d, paths, p1, p2, u)


@pytest.mark.parametrize("kwargs,reason", [
    (dict(disparities=192), "disparities must be 64 or 128"),
    (dict(paths=4), "paths must be 3 or 8"),
    (dict(p1=10, p2=10), "0 <= p1 < p2"),
    (dict(uniqueness=0.5), "uniqueness"),
])
def test_band_search_refuses_before_any_device_work(kwargs, reason):
    """nx_stereo.ex:88-131, 286-308: sgm_band/4's checks, in the library."""
    left = np.zeros((16, 80), np.uint8)
    with pytest.raises(_native.NativeError, match=reason):
        _native.sgm_band(left, left, np.zeros((16, 80), np.int32), **kwargs)


@pytest.mark.parametrize("shape,kwargs,reason", [
    ((32, 80), dict(margin=-1), "margin must be"),
    ((32, 80), dict(coarse_paths=5), "paths must be 3 or 8"),
    ((16, 16), {}, "wider than 8"),  # halves of 8x8 are too small for the 64-disparity match
    ((32, 80), dict(band=12), "band must be 8, 16, 32 or 64"),
])
def test_coarse_to_fine_refuses_before_any_device_work(shape, kwargs, reason):
    """nx_stereo.ex:164-185: sgm_coarse_to_fine/3's checks, in the library."""
    left = np.zeros(shape, np.uint8)
    with pytest.raises(_native.NativeError, match=reason):
        _native.sgm_coarse_to_fine(left, left, **kwargs)


def test_the_no_alloc_matcher_refuses_what_its_scratch_cannot_hold_before_any_device_work():
    """nxs_match_device_noalloc checks the settings and the size against its scratch before
    it touches a device (the pointers here are never read)."""
    lib = _native.load()
    fake = ctypes.c_void_p(0x1000)
    rc = lib.nxs_match_device_noalloc(fake, fake, 728, 544, 128, 8, 8, 64, 1.0526316, fake, None)
    assert rc == _native.NXS_EINVAL
    assert "more than the 67108864" in _native.last_error()
    rc = lib.nxs_match_device_noalloc(fake, fake, 64, 48, 64, 8, -1, 64, 1.0526316, fake, None)
    assert rc == _native.NXS_EINVAL and "0 <= p1" in _native.last_error()
    assert _native.NOALLOC_SCRATCH_BYTES == 64 << 20
    text = open(os.path.join(ROOT, "include", "nxs_hip.h")).read()
    assert "#define NXS_NOALLOC_SCRATCH_BYTES (64u << 20)" in text


def test_a_batch_refuses_an_unknown_mode_and_a_fused_batch_of_nine_classes_before_any_device_work():
    """nxs_batch_create checks its mode, and a fused batch's class count (one launch per
    kernel holds at most 8 size classes), before it allocates or launches anything (the
    pointers here are never read)."""
    lib = _native.load()
    fake = 0x1000
    text = open(os.path.join(ROOT, "include", "nxs_hip.h")).read()
    for name, value in (("DEFAULT", 0), ("TUNED", 1), ("FUSED", 2)):
        assert f"#define NXS_BATCH_{name} {value}" in text
        assert getattr(_native, f"BATCH_{name}") == value
    pairs = (_native.NxsPair * 9)()
    for k in range(9):
        pairs[k] = _native.NxsPair(fake, fake, fake, 64 + 8 * k, 48, 64, 8, 8, 64, 1.0526316)
    handle = ctypes.c_void_p()
    assert lib.nxs_batch_create(pairs, 9, 3, ctypes.byref(handle)) == _native.NXS_EINVAL
    assert "mode must be" in _native.last_error()
    assert lib.nxs_batch_create(pairs, 9, _native.BATCH_FUSED, ctypes.byref(handle)) != _native.NXS_OK
    assert "at most 8 size classes, got 9" in _native.last_error()
    assert not handle.value
    # A pinned plan: its fields are checked, and a fused one holds at most 8 classes too.
    for plan, reason in (((3, 0, 0), "branches must be 1, 2 or 4"),
                         ((1, -1, 0), "chunk must be 0 or more"),
                         ((2, 0, 5), "fused must be 0, 4 or 8"),
                         ((1, 0, 8, 2), "queue must be 0 or 1"),
                         ((1, 0, 0, 1), "a work queue takes a fused plan")):
        rc = lib.nxs_batch_create_planned(pairs, 9, ctypes.byref(_native.NxsBatchPlan(*plan)),
                                          ctypes.byref(handle))
        assert rc == _native.NXS_EINVAL and reason in _native.last_error(), plan
    rc = lib.nxs_batch_create_planned(pairs, 9, ctypes.byref(_native.NxsBatchPlan(2, 0, 8)),
                                      ctypes.byref(handle))
    assert rc != _native.NXS_OK and "at most 8 size classes, got 9" in _native.last_error()
    assert lib.nxs_batch_create_planned(pairs, 9, None, ctypes.byref(handle)) == _native.NXS_EINVAL
    assert not handle.value
    # One class of 2^18 pairs of 64x8: the fused census launch, a 1024-thread block a row
    # in one grid dimension, would hold 2^31 threads.
    many = 1 << 18
    pairs = (_native.NxsPair * many)()
    for k in range(many):
        pairs[k] = _native.NxsPair(fake, fake, fake, 64, 8, 64, 8, 8, 64, 1.0526316)
    assert lib.nxs_batch_create(pairs, many, _native.BATCH_FUSED, ctypes.byref(handle)) != _native.NXS_OK
    assert "would pass 2^31 threads" in _native.last_error()
    assert not handle.value


def test_the_nif_calls_only_entry_points_the_header_declares():
    """c_src/nxs_nif.c (compiled where Erlang's headers are) goes through include/nxs_hip.h
    alone: every nxs_ call it makes is declared there."""
    text = open(os.path.join(ROOT, "nx_stereo_amd", "elixir", "c_src", "nxs_nif.c")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    called = set(re.findall(r"\b(nxs_[a-z_]+)\s*\(", text))
    assert called and called <= set(declared_functions()), called - set(declared_functions())
