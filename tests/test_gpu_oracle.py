"""GPU parity against the C oracle on seeded scenes larger than the fixtures.

Each scene is stepped on the device; after every step the oracle restarts from
the device's own previous state (its broad-phase trees included), runs the pair
update over its own trees -- which fixes the new contacts and their ids
independently of the device's sort and sweep -- and computes the step itself. The
device's contact ids, manifolds, colours, impulses, bodies (fast bodies of
continuous collision included), boxes, move array, island graph and trees must
equal the oracle's bit for bit. The contact set after a step must also equal the
brute-force overlap set of the state before it.

The oracle needs the reference's own shape geometry (hulls, rounded boxes); it is
written by the reference build's fixture harness (oracle/_ref/ref_harness geom).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes
from forge2d_amd import snapshot as S
from oracle import pyoracle as O
from tests import parity_util as P
from tests.golden import make_golden

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# (scene, steps): the last two exceed the 64 KB of LDS a world may use, so their
# body states (pyramid48) and their sort-and-sweep too (container2600) run from
# global memory
SCENES = {
    "pyramid25": lambda: scenes.pyramid_world(25),
    "pyramid48": lambda: scenes.pyramid_world(48),
    "container2600": lambda: scenes.container_world(2600, 40.0, 60.0, 7003, continuous=True, chains=6,
                                                    drop_band=(0.05, 0.9)),
    "container400_ccd": lambda: scenes.container_world(400, 18.0, 40.0, 7001, continuous=True, chains=4,
                                                       drop_band=(0.3, 1.0)),
    "container300_nocd": lambda: scenes.container_world(300, 12.0, 30.0, 7002, continuous=False, chains=2,
                                                        drop_band=(0.1, 0.8)),
    # sleep enabled: islands settle and sleep, get woken, split
    "sleep_world": lambda: make_golden.sleep_world(),
    "container150_sleep": lambda: scenes.container_world(150, 10.0, 20.0, 7004, sleep=True, continuous=True,
                                                         chains=2, drop_band=(0.05, 0.5)),
}
STEPS = {"pyramid48": 12, "container2600": 12, "sleep_world": 260, "container150_sleep": 240}


def reference_geometry(spec: bytes) -> bytes:
    with tempfile.TemporaryDirectory() as td:
        sp, gp = os.path.join(td, "s.bin"), os.path.join(td, "g.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        subprocess.run([HARNESS, "geom", sp, gp, "0"], check=True, timeout=60)
        with open(gp, "rb") as f:
            return f.read()


def contacts_by_id(s):
    return {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in s.live_contacts()}


def compare(mine, o):
    assert np.array_equal(mine.pair_keys(), o.pair_keys())
    assert contacts_by_id(mine) == contacts_by_id(o)
    assert np.array_equal(mine.free_ids, o.free_ids)
    assert P.color_map(mine) == P.color_map(o)
    assert all(np.array_equal(a, c) for a, c in zip(mine.colors, o.colors)), "colour array order"
    for f in P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s"]:
        assert np.array_equal(mine.bodies[f], o.bodies[f]), f
    rep = P.manifold_report(mine, o, with_impulses=True)
    assert rep["point_count"] == rep["ids"] == rep["normals"] == rep["values"] == rep["impulses"] == 0, rep
    for f in ("aabb", "fat"):
        assert np.array_equal(mine.shapes[f], o.shapes[f]), f
    assert np.array_equal(mine.move, o.move)
    assert P.island_differences(mine, o) == []
    assert P.tree_differences(mine, o, 0.0) == []


@pytest.mark.skipif(not (O.available("f32") and os.path.exists(HARNESS)), reason="oracle or reference build missing")
@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_matches_oracle_every_step(gpu, name):
    spec = scenes.to_bytes([SCENES[name]()])
    geoms = reference_geometry(spec)
    b = F.Batch(spec)
    prev_raw = b.export(0)
    prev = S.parse_one(prev_raw)
    steps = STEPS.get(name, 40)
    for k in range(1, steps + 1):
        print(f"[oracle-parity] {name} step {k}", flush=True)
        b.step(steps=1)
        assert b.sync() == 0
        raw = b.export(0)
        mine = S.parse_one(raw)
        # the pair set by brute force over the state before the step (S/broad_phase.c:160-302)
        assert np.array_equal(O.Oracle(spec, geoms, prev_raw).pairs(), mine.pair_keys()), k
        o = S.parse_one(O.step(spec, geoms, prev_raw, prev).export())
        compare(mine, o)
        prev_raw, prev = raw, mine
