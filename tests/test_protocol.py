"""The records of native/stereo_match/src/protocol.rs:213-297, encoded and decoded."""
import struct

import numpy as np
import pytest

from nx_stereo_amd import protocol


def name(value):
    return bytes([len(value)]) + value.encode()


def test_configure_record_matches_protocol_rs():
    want = bytes([1, 7, 0, 0, 0, 64, 0, 48, 0, 128, 0, 96, 0, 2])
    want += name("median_filter") + bytes([1]) + name("kernel_size") + struct.pack("<d", 5.0)
    want += name("depth") + bytes([0])
    got = protocol.configure(7, {"width": 64, "height": 48, "frame_width": 128, "frame_height": 96},
                             [("median_filter", [("kernel_size", 5.0)]), ("depth", [])])
    assert got == want


def test_run_record_matches_protocol_rs():
    want = bytes([2, 1, 1, 0, 0, 1, 1]) + name("p1") + struct.pack("<H", 1) + struct.pack("<f", 8.0)
    want += bytes([1]) + name("left") + bytes([0, 2, 1, 0, 0, 0, 2, 0, 0, 0, 5, 6])
    got = protocol.run(257, [[("p1", np.array([8.0]))]],
                       [("left", np.array([[5, 6]], np.uint8))])
    assert got == want


def test_release_record():
    assert protocol.release(1) == bytes([3, 1, 0, 0, 0])


def test_decodes_times_then_tensors():
    record = bytes([8, 1, 10, 0, 0, 0, 1]) + name("cloud_points") + bytes([3, 0, 3, 0, 0, 0])
    out = protocol.decode(record)
    assert out.times_us == [10] and out.order == ["cloud_points"]
    assert out.tensors["cloud_points"].dtype == np.int32 and int(out.tensors["cloud_points"]) == 3


def test_decodes_configured_refused_and_no_record():
    assert protocol.decode(bytes([7])) is True
    assert protocol.decode(bytes([9]) + b"no HIP implementation of decode").reason == \
        "no HIP implementation of decode"
    assert protocol.decode(b"") is None


def test_round_trips_every_element_type():
    tensors = [("a", np.array([[1, 2], [3, 255]], np.uint8)),
               ("b", np.array([0, 1000, 65535], np.uint16)),
               ("c", np.array([[-1.0, 2.5]], np.float32)),
               ("d", np.array(-7, np.int32))]
    body = b"".join(protocol.tensor_bytes(k, v) for k, v in tensors)
    out = protocol.decode(bytes([8, 0, len(tensors)]) + body)
    for k, v in tensors:
        assert out.tensors[k].dtype == v.dtype and (out.tensors[k] == v).all()


def test_refuses_malformed_replies():
    with pytest.raises(protocol.ProtocolError):
        protocol.decode(bytes([5]))
    with pytest.raises(protocol.ProtocolError):
        protocol.decode(bytes([8, 0, 1]) + name("x") + bytes([2, 1, 4, 0, 0, 0, 0, 0]))
