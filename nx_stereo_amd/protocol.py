"""The request and reply records of the segment protocol.

Byte for byte the records NxStereo.NativeSegment writes and reads
(lib/nx_stereo/native_segment.ex:256-386) and stereo_match parses and writes
(native/stereo_match/src/protocol.rs:1-23, 107-202):

* configure: ``1, segment u32, width, height, frame_width, frame_height u16,
  member count u8``, then per member its kind (a name) and its settings
  (a count u8, then name and f64 each);
* run: ``2, segment u32, member count u8``, per member its live inputs (a
  count u8, then name, value count u16 and f32 values), then the entry
  tensors (a count u8, then each tensor);
* release: ``3, segment u32``;

and the replies: tag 7 (configured), tag 8 (member count u8, u32 us each,
output count u8, tensors) and tag 9 (refused, UTF-8 reason). A name is its
length (u8) and bytes; a tensor is its name, type u8 (0 u8, 1 u16, 2 f32,
3 s32, 4 u32: packed frames and images), rank u8, u32 dimensions and row-major elements. Little-endian.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

CONFIGURE, RUN, RELEASE = 1, 2, 3
TAG_LOG, TAG_CONFIGURED, TAG_OUTPUTS, TAG_REFUSED = 2, 7, 8, 9

# tensor.rs:25-40's types, and 4: u32, frames and images packed four pixels a word
# (packing.ex), which the segment reads as the u8 image they hold.
TYPE_CODES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2,
              np.dtype(np.int32): 3, np.dtype(np.uint32): 4}
CODE_TYPES = {v: k for k, v in TYPE_CODES.items()}


class ProtocolError(ValueError):
    """A reply that is not a record of this protocol."""


def _name(value: str) -> bytes:
    raw = value.encode("utf-8")
    if len(raw) > 255:
        raise ValueError(f"name too long: {value!r}")
    return bytes([len(raw)]) + raw


def tensor_bytes(name: str, array: np.ndarray) -> bytes:
    array = np.ascontiguousarray(array)
    code = TYPE_CODES.get(array.dtype)
    if code is None:
        raise ValueError(f"{name}: no wire type for {array.dtype}")
    head = _name(name) + bytes([code, array.ndim]) + b"".join(
        struct.pack("<I", d) for d in array.shape)
    return head + array.astype(array.dtype.newbyteorder("<"), copy=False).tobytes()


def configure(segment: int, geometry: dict, members: list[tuple[str, list[tuple[str, float]]]]) -> bytes:
    """native_segment.ex:256-276: geometry sizes, then each member's kind and settings."""
    out = [struct.pack("<BIHHHHB", CONFIGURE, segment, geometry["width"], geometry["height"],
                       geometry["frame_width"], geometry["frame_height"], len(members))]
    for kind, settings in members:
        out.append(_name(kind) + bytes([len(settings)]))
        for key, value in settings:
            out.append(_name(key) + struct.pack("<d", float(value)))
    return b"".join(out)


def run(segment: int, inputs: list[list[tuple[str, np.ndarray]]],
        entries: list[tuple[str, np.ndarray]]) -> bytes:
    """native_segment.ex:292-344: each member's params as f32, then the entry tensors."""
    out = [struct.pack("<BIB", RUN, segment, len(inputs))]
    for member in inputs:
        out.append(bytes([len(member)]))
        for key, values in member:
            values = np.asarray(values, np.float32).reshape(-1)
            out.append(_name(key) + struct.pack("<H", values.size) + values.astype("<f4").tobytes())
    out.append(bytes([len(entries)]))
    for key, array in entries:
        out.append(tensor_bytes(key, array))
    return b"".join(out)


def release(segment: int) -> bytes:
    return struct.pack("<BI", RELEASE, segment)


@dataclass
class Outputs:
    times_us: list[int]
    tensors: dict[str, np.ndarray]
    order: list[str]


@dataclass
class Refused:
    reason: str


def decode(reply: bytes):
    """A reply record: True (configured), Outputs, Refused, or None (no record)."""
    if len(reply) == 0:
        return None
    tag = reply[0]
    if tag == TAG_CONFIGURED and len(reply) == 1:
        return True
    if tag == TAG_REFUSED:
        return Refused(reply[1:].decode("utf-8", "replace"))
    if tag != TAG_OUTPUTS:
        raise ProtocolError(f"unknown reply tag {tag}")
    at = 1
    count = reply[at]
    at += 1
    times = list(struct.unpack_from(f"<{count}I", reply, at))
    at += 4 * count
    outputs = reply[at]
    at += 1
    tensors, order = {}, []
    for _ in range(outputs):
        length = reply[at]
        name = reply[at + 1:at + 1 + length].decode("utf-8")
        at += 1 + length
        code, rank = reply[at], reply[at + 1]
        at += 2
        shape = struct.unpack_from(f"<{rank}I", reply, at)
        at += 4 * rank
        dtype = CODE_TYPES.get(code)
        if dtype is None:
            raise ProtocolError(f"unknown element type {code}")
        size = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
        if at + size > len(reply):
            raise ProtocolError("the reply ends early")
        tensors[name] = np.frombuffer(reply, dtype.newbyteorder("<"), count=size // dtype.itemsize,
                                      offset=at).reshape(shape).astype(dtype)
        order.append(name)
        at += size
    if at != len(reply):
        raise ProtocolError(f"{len(reply) - at} bytes follow the reply")
    return Outputs(times, tensors, order)
