"""Talking to stereo_match_hip the way an Erlang Port does: {packet, 4} records over pipes."""
from __future__ import annotations

import os
import struct
import subprocess

PORT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nx_stereo_amd",
                    "stereo_match_hip")


def frame(record: bytes) -> bytes:
    return struct.pack(">I", len(record)) + record


def records(stream: bytes) -> list[bytes]:
    """The records of a framed byte stream (a trailing partial record is an error)."""
    out, at = [], 0
    while at < len(stream):
        (length,) = struct.unpack_from(">I", stream, at)
        at += 4
        assert at + length <= len(stream), "a record is cut short"
        out.append(stream[at:at + length])
        at += length
    return out


def converse(requests: list[bytes], args=(), timeout: float = 120.0):
    """Sends every request, closes the input, and returns (exit code, reply records, stderr)."""
    proc = subprocess.run([PORT, *args], input=b"".join(frame(r) for r in requests),
                          capture_output=True, timeout=timeout)
    return proc.returncode, records(proc.stdout), proc.stderr.decode("utf-8", "replace")
