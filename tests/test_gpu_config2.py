"""Config 2 on the GPU: one world of 20,000 mixed bodies dropped into a 200 x 200 m
segment container, continuous collision on.

* Against the reference: from the scene's initial state, the device's state after
  steps 1, 2, 3, 300, 301 and 600 hashes, field by field, to the reference's
  (tests/golden/config2_digest.json, written by tests/golden/make_golden.py from the
  reference's own snapshots of the same scene): bodies (fast bodies of continuous
  collision included), shape boxes, live contacts by id with manifolds and
  impulses, colour arrays, the free-id list, the move array and the trees. The
  later steps follow hundreds of steps with no state taken from the reference, so
  every step before them has to be bit-exact too.
* Tunnelling over 600 steps: the bodies whose centres leave the container are the
  reference's own. The reference lets one body (447) through the floor between
  steps 200 and 300, pressed by the pile above it; no other body leaves.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes
from tests import parity_util as P

HERE = os.path.dirname(os.path.abspath(__file__))
DIGESTS = os.path.join(HERE, "golden", "config2_digest.json")


def load_digests():
    with open(DIGESTS) as f:
        return json.load(f)


def config2_spec() -> bytes:
    return scenes.to_bytes(scenes.config2())


def test_config2_spec_matches_digest_fixture():
    """The scene generator still writes the scene the reference digests were taken on."""
    assert hashlib.sha256(config2_spec()).hexdigest() == load_digests()["spec_sha256"]


@pytest.mark.gpu
def test_config2_matches_reference_digests(gpu):
    from tests.golden.make_golden import config2_outside

    ref = load_digests()
    b = F.Batch(config2_spec())
    done = 0
    for want in ref["digests"]:
        print(f"[config2] stepping to {want['step']}", flush=True)
        b.step(steps=want["step"] - done)
        assert b.sync() == 0
        done = want["step"]
        snap = b.snapshot(0)
        mine = P.state_digest(snap)
        assert config2_outside(snap.bodies["px"], snap.bodies["py"]) == want["outside_container"], want["step"]
        diffs = P.digest_differences(mine, {k: v for k, v in want.items() if k != "outside_container"})
        assert diffs == [], (want["step"], diffs)
