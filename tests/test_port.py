"""stereo_match_hip without a GPU: its arguments, and the log record it answers with when the
device cannot be opened (framing.rs:134-140), framed as the Port reads it."""
import os

import pytest

from port_io import PORT, converse


pytestmark = pytest.mark.skipif(not os.path.exists(PORT), reason="stereo_match_hip is not built")


def log_of(replies):
    assert replies and replies[0][0] == 2, "a log record (tag 2) comes first"
    return replies[0][1], replies[0][2:].decode()


def test_usage_errors_are_log_records():
    code, replies, _ = converse([], args=["--bogus"])
    assert code == 2 and log_of(replies) == (2, "usage: stereo_match_hip [--threads N] [--device N], got --bogus")
    code, replies, _ = converse([], args=["--threads", "0"])
    assert code == 2 and log_of(replies) == (2, "--threads takes a positive integer, got 0")


@pytest.mark.parametrize("args", [[], ["--threads", "4"]])
def test_first_record_is_a_log(args):
    """NxStereo.Matcher passes --threads N from its options (matcher.ex:263-266)."""
    code, replies, _ = converse([], args=args)
    level, text = log_of(replies)
    if level == 2:  # no device here: the error, then the end
        assert code == 1 and "cannot open HIP device 0" in text and len(replies) == 1
    else:  # a device: the info line, then the end of the input closes it
        assert level == 0 and code == 0 and text.startswith("computing on HIP device 0")
