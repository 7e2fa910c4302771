"""The flat ABI's body, shape and joint calls against the reference, step by step.

The ops fixture (tests/golden/make_golden.py OPS_FIXTURES, written by the reference
build's "ops" mode) holds a scene and a list of API calls with the step before which
each is made: forces and torques at a point and at the centre (with and without
waking), linear and angular impulses (one beyond the world's speed limit),
transforms (a dynamic body, a static body whose proxy moves in the static tree, one
that stays inside its fat box), velocities (one on a fixed-rotation body, ignored),
mass data and mass from shapes, density, friction and restitution, gravity scale,
damping, fixed rotation on and off, bullet, putting a body to sleep, waking a jointed
chain, and a body created asleep that a call without wake leaves alone and a call
with wake wakes. The world is built through the flat ABI (tests/flat_util.py), the
calls are made through the f2d_* functions, and after every one of 40 steps the
state must equal the reference's bit for bit: bodies (and their solver sets and
sleep times), contacts by id with manifolds and impulses, colour arrays, trees,
islands and the move array.
"""
import numpy as np
import pytest

from tests import parity_util as P
from tests.conftest import GOLDEN
from tests.flat_util import FlatWorld, load_ops_fixture

pytestmark = pytest.mark.gpu


def contacts_by_id(s):
    return {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in s.live_contacts()}


def test_flat_api_calls_match_reference(gpu):
    spec, snaps, ops = load_ops_fixture("ops_bodies", GOLDEN)
    fw = FlatWorld(spec)
    ref0 = snaps[0]
    assert np.array_equal(fw.snapshot().bodies["px"], ref0.bodies["px"])
    k = 0
    for ref in snaps[1:]:
        s = ref.step
        while k < len(ops) and int(ops[k]["step"]) == s:
            fw.apply(ops[k])
            k += 1
        print(f"[flat-ops] step {s}", flush=True)
        fw.step()
        mine = fw.snapshot()
        for f in P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s", "lcx", "lcy", "inv_mass", "inv_i", "set_index",
                                  "local_index", "sleep_time", "flags"]:
            assert np.array_equal(mine.bodies[f], ref.bodies[f]), (s, f)
        assert contacts_by_id(mine) == contacts_by_id(ref), s
        rep = P.manifold_report(mine, ref, with_impulses=True)
        assert rep["point_count"] == rep["ids"] == rep["normals"] == rep["values"] == rep["impulses"] == 0, (s, rep)
        assert all(np.array_equal(a, c) for a, c in zip(mine.colors, ref.colors)), s
        for f in ("aabb", "fat", "proxy_key"):
            assert np.array_equal(mine.shapes[f], ref.shapes[f]), (s, f)
        assert np.array_equal(mine.move, ref.move), s
        assert P.tree_differences(mine, ref) == [], s
        assert P.island_differences(mine, ref) == [], s
    assert k == len(ops)
    fw.close()
