"""CPU tests: library loading, C-ABI symbols, bit-exact host scene construction and
snapshot import/export round trips against the reference's golden fixtures."""
import os
import re

import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes
from forge2d_amd import snapshot as S
from tests.conftest import FIXTURE_NAMES, ROOT, load_fixture
from tests import parity_util as P

HEADER = os.path.join(ROOT, "include", "forge2d_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    # function declarations: "<type> name(" at line start inside the extern block
    names = re.findall(r"^[A-Za-z_][A-Za-z0-9_ \*]*?\b((?:f2d|f2dg)_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    lib = F.lib()
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in F.SIGNATURES}
    assert set(names) == bound, sorted(set(names) ^ bound)


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_host_construction_matches_reference_step0(name):
    """Bodies (mass, inertia, centre, extents), shape AABBs/fat AABBs, proxy keys,
    the initial move buffer and joint colours equal the reference's bit for bit."""
    spec, snaps, _ = load_fixture(name)
    ref = snaps[0]
    assert ref.step == 0
    b = F.Batch(spec)
    mine = b.snapshot(0)
    for f in ["type", "px", "py", "qc", "qs", "cx", "cy", "lcx", "lcy", "inv_mass", "inv_i", "min_ext",
              "max_ext", "vx", "vy", "w"]:
        assert np.array_equal(mine.bodies[f], ref.bodies[f]), f
    assert np.array_equal(mine.shapes["fat"], ref.shapes["fat"])
    assert np.array_equal(mine.shapes["aabb"], ref.shapes["aabb"])
    assert np.array_equal(mine.shapes["proxy_key"], ref.shapes["proxy_key"])
    assert np.array_equal(mine.shapes["body"], ref.shapes["body"])
    assert np.array_equal(mine.move, ref.move)
    assert np.array_equal(mine.joints["color"], ref.joints["color"])
    # solver sets and islands at creation: one island per non-static body, joints
    # linked and their islands merged at once (S/body.c:84-97, S/joint.c:289-309)
    assert P.island_differences(mine, ref) == []
    # broad-phase trees built by inserting each proxy with rotations (S/dynamic_tree.c:625-795)
    assert P.tree_differences(mine, ref) == []


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_snapshot_import_export_roundtrip(name):
    """Importing a reference state and exporting it again reproduces the state:
    bodies, contacts by id with manifolds, colour lists, overflow order, move
    buffer and the LIFO contact-id free list."""
    spec, snaps, raws = load_fixture(name)
    b = F.Batch(spec)
    for ref, raw in zip(snaps[1:], raws[1:]):
        b.import_state(0, raw)
        back = b.snapshot(0)
        assert P.body_mismatch_count(back, ref, P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s"]) == 0
        assert np.array_equal(back.live_contacts()["id"], ref.live_contacts()["id"])
        rep = P.manifold_report(back, ref, with_impulses=True)
        assert not rep["only_mine"] and not rep["only_ref"]
        assert rep["point_count"] == rep["ids"] == rep["values"] == rep["impulses"] == 0
        assert P.color_map(back) == P.color_map(ref)
        assert all(np.array_equal(a, c) for a, c in zip(back.colors, ref.colors)), "colour array order"
        assert np.array_equal(back.move, ref.move)
        assert np.array_equal(back.free_ids, ref.free_ids)
        assert P.island_differences(back, ref) == []
        assert P.tree_differences(back, ref) == []


def test_scene_generators_are_deterministic():
    a = scenes.to_bytes(scenes.config3(worlds=8))
    b = scenes.to_bytes(scenes.config3(worlds=8))
    assert a == b
    worlds = scenes.parse(a)
    assert len(worlds) == 8
    rows = [int(round((np.sqrt(8 * (h["body_count"] - 1) + 1) - 1) / 2)) for h, _, _, _ in worlds]
    assert all(5 <= r <= 30 for r in rows)
    c4 = scenes.to_bytes(scenes.config4(worlds=6))
    assert scenes.to_bytes(scenes.config4(worlds=6)) == c4


def test_config_scenes_build_in_the_library():
    """The generated configuration scenes (small instances) are accepted by the
    host builder: every polygon hull is valid and joints reference real bodies."""
    for worlds in (scenes.config1(), scenes.config3(worlds=4), scenes.config4(worlds=4),
                   [scenes.container_world(300, 20.0, 40.0, 5, chains=4)]):
        b = F.Batch(worlds)
        assert b.world_count == len(worlds)
        for i, w in enumerate(worlds):
            h, bodies, shapes, joints = w.arrays()
            assert F.lib().f2dg_body_count(b.h, i) == len(bodies)
            assert F.lib().f2dg_shape_count(b.h, i) == len(shapes)


def test_degenerate_hull_is_rejected():
    w = scenes.World(scenes.world_header())
    g = w.add_body(scenes.DYNAMIC, 0.0, 0.0)
    w.add_polygon(g, [(0.0, 0.0), (1.0, 0.0), (2.0, 0.0)])  # collinear
    with pytest.raises(ValueError):
        F.Batch([w])


def test_post_collide_composition_matches_persisting_impulses():
    """warm_start_matched copies impulses by feature id: for a settled pyramid the
    manifolds of consecutive steps keep their ids, so every point persists."""
    _, snaps, _ = load_fixture("pyramid12")
    prev, cur = snaps[2], snaps[3]
    m = S.warm_start_matched(prev, cur)
    live = m[m["id"] >= 0]
    old = prev.contact_by_pair()
    for rec in live:
        o = old[(int(rec["shape_a"]), int(rec["shape_b"]))]
        for p in range(int(rec["point_count"])):
            assert rec["pts"][p]["persisted"] == 1
            assert rec["pts"][p]["ni"] == o["pts"][p]["ni"]
