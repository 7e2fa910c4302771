"""GPU parity against the reference's own step (golden fixtures written by the
reference Box2D through oracle/ref_harness.c).

For every pair of consecutive dumped steps (k-1, k) of every fixture:
  * full step: import the reference state after step k-1, run one device step and
    compare with the reference after step k: the contact pair set and every
    manifold (point count, feature ids, anchors, separation, normal) bit-exact;
    new contacts get the reference's ids (the pair update follows the broad-phase
    trees, which are kept and compared node for node); colours, the overflow order
    and the free-id list equal; body position, rotation and velocity within 1e-5
    relative (floor 1: absolute below unit magnitude), and in fact bit for bit:
    continuous collision visits its candidates in the trees' query order.
  * solver + finalize: the post-collide state of step k (colours, manifolds,
    impulses matched from k-1, bodies from k-1, trees rebuilt), as the oracle's
    pair update and collide compute it from the reference's step k-1, stepped
    through the solver stages and finalize reproduces the reference's bodies,
    stored impulses, boxes and next move buffer of step k, bit for bit.
The ccd fixtures fire fast bodies and bullets at thin static geometry, so these
cases also cover the time-of-impact sweeps of continuous collision. The sleep
fixture (every step 0-260) covers islands: links, merges, splits, islands put to
sleep and woken, compared exactly (island records, id pools, awake orders, links)
on every step.
"""
import os

import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import snapshot as S
from forge2d_amd import scenes
from oracle import pyoracle as O
from tests import parity_util as P
from tests.conftest import FIXTURE_NAMES, GOLDEN, load_fixture

pytestmark = pytest.mark.gpu
TOL = 1e-5


def consecutive_pairs(name):
    spec, snaps, raws = load_fixture(name)
    out = []
    for i in range(1, len(snaps)):
        if snaps[i].step == snaps[i - 1].step + 1:
            out.append((name, i))
    return out


CASES = [c for n in FIXTURE_NAMES for c in consecutive_pairs(n)]


def contacts_by_id(s):
    return {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in s.live_contacts()}


@pytest.mark.parametrize("name,i", CASES, ids=[f"{n}-{load_fixture(n)[1][i].step}" for n, i in CASES])
def test_full_step_from_reference_state(gpu, name, i):
    spec, snaps, raws = load_fixture(name)
    prev, ref = snaps[i - 1], snaps[i]
    print(f"[parity] full step {name} {prev.step}->{ref.step}", flush=True)
    b = F.Batch(spec)
    b.import_state(0, raws[i - 1])
    b.step(steps=1)
    assert b.sync() == 0
    mine = b.snapshot(0)
    assert np.array_equal(np.sort(mine.pair_keys()), np.sort(ref.pair_keys())), "contact pair set differs"
    # new contacts take the reference's ids: the pair update follows its trees
    assert contacts_by_id(mine) == contacts_by_id(ref), "contact ids differ"
    assert np.array_equal(mine.free_ids, ref.free_ids)
    rep = P.manifold_report(mine, ref)
    assert rep["point_count"] == 0 and rep["ids"] == 0, rep
    assert rep["values"] == 0 and rep["normals"] == 0, rep
    assert P.coloring_violations(mine) == 0
    assert P.color_map(mine) == P.color_map(ref)
    assert all(np.array_equal(a, c) for a, c in zip(mine.colors, ref.colors)), "colour array order"
    errs = P.body_errors(mine, ref)
    assert max(errs.values()) <= TOL, errs
    # in practice bit for bit, fast bodies of continuous collision included
    assert P.body_mismatch_count(mine, ref, P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s"]) == 0
    rep = P.manifold_report(mine, ref, with_impulses=True)
    assert rep["impulses"] == 0, rep
    assert P.island_differences(mine, ref) == []
    assert P.tree_differences(mine, ref, 0.0) == []
    assert np.array_equal(mine.move, ref.move)


def sleep_enabled(spec: bytes) -> bool:
    return bool(int(scenes.parse(spec)[0][0]["flags"]) & scenes.WF_SLEEP)


def fixture_geoms(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))["geoms"].tobytes()


SOLVE_CASES = [c for c in CASES if load_fixture(c[0])[1][c[1]].live_contacts().size > 0]


@pytest.mark.parametrize("name,i", SOLVE_CASES, ids=[f"{n}-{load_fixture(n)[1][i].step}" for n, i in SOLVE_CASES])
def test_solver_and_finalize_from_post_collide_state(gpu, name, i):
    spec, snaps, raws = load_fixture(name)
    prev, ref = snaps[i - 1], snaps[i]
    print(f"[parity] solve {name} {prev.step}->{ref.step}", flush=True)
    # the post-collide state as the oracle's pair update and collide compute it
    # (pinned bit-exact to the reference): contacts, manifolds and matched impulses,
    # the island graph, and the trees rebuilt at the start of the step, whose query
    # order continuous collision follows
    o = O.Oracle(spec, fixture_geoms(name), raws[i - 1])
    o.update_pairs()
    o.collide()
    post = o.export()
    ref_post = S.post_collide(prev, ref)
    assert np.array_equal(S.parse_one(post).pair_keys(), ref_post.pair_keys())
    b = F.Batch(spec)
    b.import_state(0, post)
    b.step(steps=1, phases=F.PH_SOLVE | F.PH_FINALIZE)
    assert b.sync() == 0
    mine = b.snapshot(0)
    errs = P.body_errors(mine, ref, P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s"])
    assert max(errs.values()) <= TOL, errs
    assert P.body_mismatch_count(mine, ref, P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s"]) == 0
    rep = P.manifold_report(mine, ref, value_fields=(), with_impulses=True)
    assert rep["impulses"] == 0, rep
    for f in ("aabb", "fat"):
        assert np.array_equal(mine.shapes[f], ref.shapes[f]), f
    assert np.array_equal(mine.move, ref.move)
    for f in ("linear_impulse_x", "linear_impulse_y", "lower_impulse", "upper_impulse"):
        assert np.array_equal(mine.joints[f], ref.joints[f]), f
    if sleep_enabled(spec):
        assert P.island_differences(mine, ref) == []


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_first_step_pairs_and_manifolds(gpu, name):
    """Stepping the freshly built scene: the created pair set and every manifold
    equal the reference's step-1 result bit for bit."""
    spec, snaps, _ = load_fixture(name)
    if snaps[1].step != 1:
        pytest.skip("fixture has no step 1")
    b = F.Batch(spec)
    b.step(steps=1)
    assert b.sync() == 0
    mine = b.snapshot(0)
    ref = snaps[1]
    assert np.array_equal(np.sort(mine.pair_keys()), np.sort(ref.pair_keys()))
    assert contacts_by_id(mine) == contacts_by_id(ref)
    rep = P.manifold_report(mine, ref)
    assert rep["point_count"] == rep["ids"] == rep["values"] == rep["normals"] == 0, rep
    assert P.coloring_violations(mine) == 0
    assert P.color_map(mine) == P.color_map(ref)
    assert P.tree_differences(mine, ref, TOL) == []
