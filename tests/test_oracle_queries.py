"""Pins of the oracle's world queries (oracle/b2oracle.c, b2o_query) against the
reference's own b2World_CastRayClosest, b2World_CastRay, b2World_OverlapAABB and
b2World_Explode (S/world.c:2050-2308, 2640-2747), run by the reference build on the
fixture worlds (tests/golden/make_golden.py, harness mode "query").

"queries": every shape type on static, kinematic and dynamic bodies, rounded boxes
(ray cast by the shape cast, S/distance.c:608-700), a two-shape body, categories 1,
2, 4 and one in the high word with masks that leave category 2 out; 160 rays (from
inside bodies, of zero length, every filter), 40 AABB boxes, five explosions.
"queries_sleep": the sleep scene after 200 steps; an explosion wakes a sleeping
pyramid (one island). "queries_wake2": one explosion wakes two boxes sleeping apart
(two islands).

Every result must agree bit for bit and in the reference's order: rays walk each
tree nearer child first with the segment shortened by accepted hits, overlaps and
explosions walk it depth first (child2 popped first). After the explosions every
body, the island graph, the solver sets (the order explosions wake sleeping sets
in) and the colouring must equal the reference's.
"""
import os

import numpy as np
import pytest

from forge2d_amd import QUERY_DTYPE, snapshot as S
from oracle import pyoracle as O
from tests import parity_util as P
from tests.conftest import GOLDEN, load_fixture
from tests.golden.make_golden import parse_results

pytestmark = pytest.mark.skipif(not O.available("f32"), reason="oracle not built (make -C oracle port)")


def query_fixture(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    spec, snaps, raws = load_fixture(name)
    q = np.frombuffer(z["queries"].tobytes(), QUERY_DTYPE)
    return spec, z["geoms"].tobytes(), snaps, raws, q, parse_results(z["results"].tobytes(), len(q))


@pytest.mark.parametrize("name", P.QUERY_FIXTURES)
def test_oracle_queries_match_reference(name):
    spec, geoms, snaps, raws, q, ref = query_fixture(name)
    o = O.Oracle(spec, geoms, raws[0])
    counts, hits = o.query(q, cap=256)
    assert P.query_mismatches(q, ref, counts, hits) == []
    check_after_explosions(name, S.parse_one(o.export()), snaps[1])


def check_after_explosions(name, after, ref):
    for f in P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s", "set_index", "local_index", "sleep_time"]:
        assert np.array_equal(after.bodies[f], ref.bodies[f]), f
    assert P.island_differences(after, ref) == []
    assert P.color_map(after) == P.color_map(ref)
    assert all(np.array_equal(a, c) for a, c in zip(after.colors, ref.colors)), "colour array order"


@pytest.mark.parametrize("name", P.QUERY_FIXTURES)
def test_query_fixture_exercises_every_case(name):
    """The fixture holds hits of every kind, initial overlaps, and explosions that
    move (and in the sleep scene wake) bodies, so a dropped case cannot pass."""
    spec, geoms, snaps, raws, q, ref = query_fixture(name)
    for kind in (1, 2, 3):
        assert any(len(ref[i]) for i in range(len(q)) if q[i]["kind"] == kind), kind
    assert any(len(ref[i]) > 1 for i in range(len(q)) if q[i]["kind"] == 2)
    assert any((ref[i]["fraction"] == 0).any() for i in range(len(q)) if q[i]["kind"] == 2)
    b0, b1 = snaps[0].bodies, snaps[1].bodies
    assert np.sum(b0["vx"] != b1["vx"]) >= (2 if name == "queries_wake2" else 6)
    if name != "queries":
        woken = (b0["set_index"] >= 3) & (b1["set_index"] == 2)
        assert len(set(b0["set_index"][woken])) == (2 if name == "queries_wake2" else 1)
    else:
        types = np.frombuffer(geoms, np.int32).reshape(-1, 37)[:, 0]
        radius = np.frombuffer(geoms, np.float32).reshape(-1, 37)[:, 2]
        hit_shapes = np.concatenate([r["shape"] for r in ref if len(r)])
        assert set(types[hit_shapes]) == {0, 1, 2, 3}
        assert ((types[hit_shapes] == 3) & (radius[hit_shapes] > 0)).any()


def test_fixture_orders_are_not_sorted_orders():
    """The reference's every-hit and overlap results come in tree-walk order, which is
    neither shape-id nor fraction order for some queries, so the order is pinned."""
    spec, geoms, snaps, raws, q, ref = query_fixture("queries")
    rays = [ref[i] for i in range(len(q)) if q[i]["kind"] == 2 and len(ref[i]) > 1]
    boxes = [ref[i] for i in range(len(q)) if q[i]["kind"] == 3 and len(ref[i]) > 1]
    assert any(not np.all(np.diff(r["fraction"]) >= 0) for r in rays)
    assert any(not np.all(np.diff(r["shape"]) > 0) for r in rays)
    assert any(not np.all(np.diff(r["shape"]) > 0) for r in boxes)


def test_planted_mistake_is_caught():
    """A closest-hit fraction changed in its last bit, a dropped overlap, or two
    overlaps swapped fails the comparison."""
    spec, geoms, snaps, raws, q, ref = query_fixture("queries")
    o = O.Oracle(spec, geoms, raws[0])
    counts, hits = o.query(q, cap=256)
    i = next(k for k in range(len(q)) if q[k]["kind"] == 1 and counts[k] == 1)
    h = hits.copy()
    h[i][0]["fraction"] = np.nextafter(h[i][0]["fraction"], np.float32(2))
    assert P.query_mismatches(q, ref, counts, h) == [i]
    j = next(k for k in range(len(q)) if q[k]["kind"] == 3 and counts[k] > 1)
    c = counts.copy()
    c[j] -= 1
    assert P.query_mismatches(q, ref, c, hits) == [j]
    h = hits.copy()
    h[j][[0, 1]] = h[j][[1, 0]]
    assert P.query_mismatches(q, ref, counts, h) == [j]
