"""stereo_match_hip on the GPU, driven as NxStereo.Matcher drives stereo_match over a Port.

Two segments of different sizes are configured and run interleaved; a malformed request, a
run of an unknown segment and a run after a release are refused; every reply comes in the
order of the requests (main.rs:7-9) and every output is the oracle's, bit for bit. The test
process itself never opens the device: the executable does.
"""
import os

import numpy as np
import pytest

from nx_stereo_amd import protocol, scenes, stages
from nx_stereo_amd.segment import HipSegment
from oracle import oracle
from port_io import PORT, converse

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(PORT), reason="stereo_match_hip is not built")]


def geometry(w, h):
    return {"width": w, "height": h, "frame_width": w, "frame_height": h, "focal_length_px": 873.0,
            "baseline_m": 0.2, "principal_point": (w / 2, h / 2)}


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_port_conversation_matches_the_oracle():
    a = HipSegment(None, [("disparity", stages.Disparity(num_disparities=64, paths=8)),
                          ("speckle_filter", stages.SpeckleFilter())], geometry(97, 31))
    b = HipSegment(None, [("disparity", stages.Disparity(num_disparities=128, paths=8)),
                          ("temporal_filter", stages.TemporalFilter(alpha=0.5, delta=2.0, hold=1)),
                          ("quality", stages.Quality())], geometry(160, 48))
    pairs_a = [scenes.planar_pair(97, 31, seed=s, low=4, high=50)[:2] for s in (1, 2)]
    pairs_b = [scenes.planar_pair(160, 48, seed=s, low=4, high=100)[:2] for s in (3, 4)]
    run = lambda seg, pair: seg.run_record({"left": pair[0], "right": pair[1]})
    requests = [a.configure_record(), b.configure_record(),
                run(a, pairs_a[0]), run(b, pairs_b[0]), run(a, pairs_a[1]),
                b"\x02\x01",                                  # cut short
                protocol.run(999, [], []),                    # never configured
                protocol.release(a.id), run(a, pairs_a[0]),   # released: refused
                run(b, pairs_b[1])]
    code, replies, err = converse(requests)
    assert code == 0, err
    log, *replies = replies
    assert log[0] == 2 and log[1] == 0, log
    decoded = [protocol.decode(r) for r in replies]
    assert decoded[0] is True and decoded[1] is True
    kinds = ["outputs" if isinstance(d, protocol.Outputs) else "refused" for d in decoded[2:]]
    assert kinds == ["outputs"] * 3 + ["refused"] * 3 + ["outputs"], kinds
    assert "segment 999 is not configured" in decoded[6].reason
    assert f"segment {a.id} is not configured" in decoded[7].reason

    for d, pair in ((decoded[2], pairs_a[0]), (decoded[4], pairs_a[1])):
        want = oracle.despeckle(oracle.sgm(pair[0], pair[1], 64, 8), 12, 1.0)
        assert same(d.tensors["disparity"], want)
    temporal, quality = oracle.Temporal((48, 160)), oracle.Quality((48, 160))
    for d, pair in ((decoded[3], pairs_b[0]), (decoded[8], pairs_b[1])):
        want = temporal.filter(oracle.sgm(pair[0], pair[1], 128, 8), 0.5, 2.0, 1)
        assert same(d.tensors["disparity"], want)
        assert same(d.tensors["quality_sums"], quality.sums(want))
