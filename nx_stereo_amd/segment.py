"""HipSegment: NxStereo.Segment's contract over the in-process C ABI.

A host-side mirror of what the Elixir `NxStereo.HipSegment` does (its NIF
binds the same C calls): the callbacks of lib/nx_stereo/segment.ex:63-137
with NxStereo.NativeSegment's semantics (lib/nx_stereo/native_segment.ex).

* ``build`` refuses members without a native kind or computed in Nx, and
  never touches the device: the segment is configured with its first pair.
* ``send_request`` sends a pair without waiting and returns its request id;
  pairs of one segment are computed in the order sent.
* ``await_`` returns ``("ok", outputs, segment, done_ns)``, ``("error",
  reason)`` for a pair that failed, ``("stop", reason)`` when the
  configuration was refused, or ``"timeout"``.
* The device keeps what stages carry: ``merge_carry`` returns the current
  segment; ``reload`` configures again, which restarts the carries.
"""
from __future__ import annotations

import ctypes
import itertools
import time

import numpy as np

from . import _native, protocol, stages

_segment_ids = itertools.count(1)


class Context:
    """An nxs_ctx: the device, its streams and memory, and the thread computing requests."""

    def __init__(self, device: int = 0):
        lib = _native.load()
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.nxs_open(device, ctypes.byref(handle), err, len(err))
        if rc != _native.NXS_OK:
            raise _native.NativeError(f"nxs_open({device}) failed: {err.value.decode()}")
        self.lib, self.handle = lib, handle

    def close(self) -> None:
        if self.handle:
            self.lib.nxs_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, record: bytes) -> int:
        buf = (ctypes.c_uint8 * len(record)).from_buffer_copy(record)
        ticket = self.lib.nxs_submit(self.handle, buf, len(record), _native.DONE_FN(), None)
        if ticket < 0:
            raise _native.NativeError(f"nxs_submit failed: {_native.last_error()}")
        return ticket

    def wait(self, ticket: int, timeout_ms: int = -1) -> bool:
        rc = self.lib.nxs_wait(self.handle, ticket, timeout_ms)
        if rc == _native.NXS_TIMEOUT:
            return False
        if rc != _native.NXS_OK:
            raise _native.NativeError(f"nxs_wait failed: {_native.last_error()}")
        return True

    def reply(self, ticket: int) -> bytes:
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t()
        rc = self.lib.nxs_reply(self.handle, ticket, ctypes.byref(ptr), ctypes.byref(size))
        if rc != _native.NXS_OK:
            raise _native.NativeError(f"nxs_reply failed: {_native.last_error()}")
        data = ctypes.string_at(ptr, size.value) if size.value else b""
        self.lib.nxs_release_reply(self.handle, ticket)
        return data

    def coalescing(self, max_pairs: int = -1):
        """nxs_coalescing: sets how many served runs of one matcher size class may share a
        launch (0 or 1: none; negative leaves it), and returns (launches, pairs) so far."""
        launches, pairs = ctypes.c_int64(), ctypes.c_int64()
        rc = self.lib.nxs_coalescing(self.handle, max_pairs, ctypes.byref(launches),
                                     ctypes.byref(pairs))
        if rc != _native.NXS_OK:
            raise ValueError(_native.last_error())
        return launches.value, pairs.value

    def request(self, record: bytes, timeout_ms: int = -1):
        ticket = self.submit(record)
        if not self.wait(ticket, timeout_ms):
            raise TimeoutError(f"ticket {ticket}")
        return protocol.decode(self.reply(ticket))

    def run_device(self, record: bytes, entries=(), stream: int = 0, max_outputs: int = 16):
        """nxs_run_device: entries are (name, type code, shape, device address)."""
        n = len(entries)
        arr = (NxsTensor * max(n, 1))()
        names = []
        for i, (name, code, shape, ptr) in enumerate(entries):
            names.append(name.encode())
            arr[i].name = names[-1]
            arr[i].type = code
            arr[i].rank = len(shape)
            for d, v in enumerate(shape):
                arr[i].shape[d] = v
            arr[i].data = ptr
        outs = (NxsTensor * max_outputs)()
        count = ctypes.c_int()
        buf = (ctypes.c_uint8 * len(record)).from_buffer_copy(record)
        rc = self.lib.nxs_run_device(self.handle, buf, len(record), arr, n, outs, max_outputs,
                                     ctypes.byref(count), ctypes.c_void_p(stream or None))
        if rc != _native.NXS_OK:
            raise _native.NativeError(f"nxs_run_device failed: {_native.last_error()}")
        return [(outs[k].name.decode(), outs[k].type, tuple(outs[k].shape[:outs[k].rank]),
                 outs[k].data) for k in range(min(count.value, max_outputs))]


NxsTensor = _native.NxsTensor

PROFILED_KERNELS = ("census_cost", "path", "select", "check_median", "band_cost", "band_path",
                    "band_select", "shrink")


def profile_start(capacity: int = 4096) -> None:
    lib = _native.load()
    if lib.nxs_profile_start(capacity) != _native.NXS_OK:
        raise _native.NativeError(_native.last_error())


def profile_stop() -> dict:
    lib = _native.load()
    n = len(PROFILED_KERNELS)
    ms, count = (ctypes.c_double * n)(), (ctypes.c_int * n)()
    if lib.nxs_profile_stop(ms, count, n) != _native.NXS_OK:
        raise _native.NativeError(_native.last_error())
    return {k: (ms[i], count[i]) for i, k in enumerate(PROFILED_KERNELS)}


class HipSegment:
    """A run of stages computed by libnxs_hip on one GPU (native_segment.ex:45-246)."""

    def __init__(self, context: Context, members, geometry):
        self.context = context
        self.members = list(members)  # (name, stage) pairs
        self.geometry = geometry
        self.id = next(_segment_ids)
        self.version = 0
        self.configured_version = None
        self.times_us: list[int] = []
        self._refresh_inputs()

    # -- Segment callbacks ---------------------------------------------------------------

    @staticmethod
    def compute():
        return {"runs_on": "gpu", "implementation": "hip", "backend": "rocm"}

    @staticmethod
    def depth() -> int:
        """Two pairs: one computing while the next is queued behind it."""
        return 2

    @staticmethod
    def refusal(stage) -> str | None:
        name = type(stage).__name__
        if not getattr(stage, "KIND", ""):
            return f"{name} has no native implementation"
        if stage.compute() == "nx":
            return f"{name} is computed in Nx, as its settings ask"
        return None

    @classmethod
    def accepts(cls, stage, compute: str) -> bool:
        return compute == "native" and cls.refusal(stage) is None

    @classmethod
    def build(cls, members, geometry, context: Context | None = None):
        """('ok', segment) or ('error', reason); the device is configured on the first pair."""
        for _, stage in members:
            reason = cls.refusal(stage)
            if reason:
                return ("error", reason)
        return ("ok", cls(context or Context(), members, geometry))

    def reload(self, members, geometry):
        self.members, self.geometry = list(members), geometry
        self.version += 1
        self._refresh_inputs()
        return self

    def put_member_state(self, index: int, state: dict):
        self.members[index][1].state = state
        self._refresh_inputs()
        return self

    def merge_carry(self, other):
        return self

    def member_times(self):
        return [(name, us * 1000) for (name, _), us in zip(self.members, self.times_us)]

    def collect(self, outputs, timeout=None):
        return ("ok", outputs)

    # -- records -------------------------------------------------------------------------

    def configure_record(self) -> bytes:
        members = [(stage.KIND, stages.build_settings(stage)) for _, stage in self.members]
        return protocol.configure(self.id, self.geometry, members)

    def _refresh_inputs(self):
        self.inputs = [stages.flatten(stage.params(self.geometry)) for _, stage in self.members]

    def run_record(self, entries: dict) -> bytes:
        return protocol.run(self.id, self.inputs, list(entries.items()))

    def send_request(self, entries: dict):
        """The pair's request id; configures first when the segment's version changed."""
        configure_ticket = None
        if self.configured_version != self.version:
            configure_ticket = self.context.submit(self.configure_record())
            self.configured_version = self.version
        return (configure_ticket, self.context.submit(self.run_record(entries)))

    def await_(self, request_id, timeout_ms: int = -1):
        configure_ticket, ticket = request_id
        if configure_ticket is not None:
            if not self.context.wait(configure_ticket, timeout_ms):
                return "timeout"
            reply = protocol.decode(self.context.reply(configure_ticket))
            if isinstance(reply, protocol.Refused):
                self.context.wait(ticket, -1)
                self.context.reply(ticket)
                self.configured_version = None
                return ("stop", reply.reason)
        if not self.context.wait(ticket, timeout_ms):
            return "timeout"
        reply = protocol.decode(self.context.reply(ticket))
        if isinstance(reply, protocol.Refused):
            return ("error", reply.reason)
        self.times_us = reply.times_us
        return ("ok", reply.tensors, self, time.monotonic_ns())

    def release(self) -> None:
        self.context.request(protocol.release(self.id))

    def run(self, entries: dict, timeout_ms: int = -1):
        """Segment.run/4 (segment.ex:154-166): the pair's outputs merged over its entries."""
        result = self.await_(self.send_request(entries), timeout_ms)
        if isinstance(result, tuple) and result[0] == "ok":
            merged = dict(entries)
            merged.update(result[1])
            return ("ok", merged, result[2])
        return result


def as_entries(**arrays) -> dict:
    return {k: np.ascontiguousarray(v) for k, v in arrays.items()}
