"""stereo_match_hip without a GPU: its arguments, and the log record it answers with when the
device cannot be opened (framing.rs:134-140), framed as the Port reads it."""
import os

import pytest

from port_io import PORT, converse


pytestmark = pytest.mark.skipif(not os.path.exists(PORT), reason="stereo_match_hip is not built")


def test_usage_error():
    code, replies, err = converse([], args=["--threads", "4"])
    assert code == 2 and replies == [] and "usage: stereo_match_hip" in err


def test_first_record_is_a_log():
    code, replies, _ = converse([])
    assert replies, "no record at all"
    tag, level, text = replies[0][0], replies[0][1], replies[0][2:].decode()
    assert tag == 2, "a log record (tag 2)"
    if level == 2:  # no device here: the error, then the end
        assert code == 1 and "cannot open HIP device 0" in text and len(replies) == 1
    else:  # a device: the info line, then the end of the input closes it
        assert level == 0 and code == 0 and text == "computing on HIP device 0"
