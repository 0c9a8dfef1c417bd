"""The oracle reproduces the committed frame digests (tests/golden)."""
import json
import os

from golden.make_fixtures import frames

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_frames_match_committed_digests():
    want = json.load(open(os.path.join(HERE, "golden", "oracle_frames.json")))
    got = frames()
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g == w, (g["kind"], g["args"])
