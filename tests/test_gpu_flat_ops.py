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
import ctypes

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


JOINT_FIELDS = ["color", "local_index", "body_a", "body_b", "enable_limit", "enable_motor", "enable_spring",
                "linear_impulse_x", "linear_impulse_y", "spring_impulse", "motor_impulse", "lower_impulse",
                "upper_impulse", "ref_angle", "reserved"]


def replay_joint_fixture(name, joints):
    spec, snaps, ops = load_ops_fixture(name, GOLDEN)
    fw = FlatWorld(spec)
    k = 0
    for ref in snaps[1:]:
        s = ref.step
        while k < len(ops) and int(ops[k]["step"]) == s:
            fw.apply(ops[k])
            k += 1
        print(f"[{name}] step {s}", flush=True)
        fw.step()
        mine = fw.snapshot()
        for f in P.BODY_FIELDS + ["set_index", "local_index"]:
            assert np.array_equal(mine.bodies[f], ref.bodies[f]), (s, f)
        assert len(mine.joints) == len(ref.joints) == joints, s
        for f in JOINT_FIELDS:
            assert np.array_equal(mine.joints[f], ref.joints[f]), (s, f)
        assert contacts_by_id(mine) == contacts_by_id(ref), s
        rep = P.manifold_report(mine, ref, with_impulses=True)
        assert rep["point_count"] == rep["ids"] == rep["normals"] == rep["values"] == rep["impulses"] == 0, (s, rep)
        assert all(np.array_equal(a, c) for a, c in zip(mine.colors, ref.colors)), s
        assert P.island_differences(mine, ref) == [], s
    assert k == len(ops)
    return fw, snaps


def test_weld_and_wheel_joints_match_reference(gpu):
    """Weld and wheel joints (S/weld_joint.c, S/wheel_joint.c) made through
    f2d_create_weld_joint / f2d_create_wheel_joint and their setters (fixture
    ops_weld_wheel: a rigid weld to the ground, a weld soft in both senses between
    two boxes, a weld with a reference angle and an angular spring only, a two-wheel
    chassis whose rear wheel has limit, motor and spring and whose front wheel takes
    an unnormalized axis and collideConnected; later a reversed motor, limits moved
    so that the lower and then the upper one acts, a spring and a limit turned off
    and a stiffer weld): after every one of 60 steps bodies, contacts, colours,
    islands and every joint's impulses equal the reference's bit for bit."""
    fw, snaps = replay_joint_fixture("ops_weld_wheel", 5)
    import forge2d_amd as F
    L = F.lib()
    assert [L.f2d_joint_get_type(*j) for j in fw.joints] == [6, 6, 7, 7, 6]
    assert L.f2d_weld_joint_get_linear_hertz(*fw.joints[1]) == 20.0
    assert L.f2d_wheel_joint_is_limit_enabled(*fw.joints[2]) == 0
    inv_h = np.float32(4.0) * (np.float32(1.0) / np.float32(1.0 / 60.0))
    assert L.f2d_wheel_joint_get_motor_torque(*fw.joints[2]) == inv_h * snaps[-1].joints["motor_impulse"][2]
    assert L.f2d_joint_get_constraint_torque(*fw.joints[0]) == inv_h * snaps[-1].joints["spring_impulse"][0]
    fw.close()


def test_revolute_joint_calls_match_reference(gpu):
    """The f2d_revolute_joint_* calls (fixture ops_revolute on the revolute chain of
    ops_bodies: a motor turned on and later reversed, limits moved and turned off, a
    spring with a target angle turned on and off): after every one of 40 steps the
    state and every joint's impulses equal the reference's bit for bit."""
    fw, snaps = replay_joint_fixture("ops_revolute", 3)
    import forge2d_amd as F
    L = F.lib()
    j0 = fw.joints[0]
    assert L.f2d_revolute_joint_get_motor_speed(*j0) == -3.0
    assert (L.f2d_revolute_joint_get_lower_limit(*j0), L.f2d_revolute_joint_get_upper_limit(*j0)) \
        == (np.float32(-0.3), np.float32(0.3))
    assert L.f2d_revolute_joint_get_target_angle(*fw.joints[1]) == np.float32(0.25)
    assert L.f2d_revolute_joint_is_spring_enabled(*fw.joints[1]) == 0
    assert L.f2d_revolute_joint_is_spring_enabled(*fw.joints[2]) == 1
    fw.close()


def test_motor_mouse_filter_joints_match_reference(gpu):
    """Motor, mouse and filter joints (S/motor_joint.c, S/mouse_joint.c, S/joint.c:480-505)
    made through f2d_create_motor_joint / f2d_create_mouse_joint / f2d_create_filter_joint
    and the f2d_motor_joint_* / f2d_mouse_joint_* calls (fixture ops_motor_mouse: a motor
    from the ground held at its force limit, a motor on a moving carrier at its torque
    limit and later with zero force, a stiff mouse joint dragged to new targets, a weak
    one held at its force limit, a filter joint between two overlapping boxes): after
    every one of 60 steps bodies, contacts, colours, islands and every joint's impulses
    equal the reference's bit for bit."""
    fw, snaps = replay_joint_fixture("ops_motor_mouse", 5)
    import forge2d_amd as F
    L = F.lib()
    assert [L.f2d_joint_get_type(*j) for j in fw.joints] == [2, 2, 3, 3, 1]
    inv_h = np.float32(4.0) * (np.float32(1.0) / np.float32(1.0 / 60.0))
    out = (ctypes.c_float * 2)()
    L.f2d_joint_get_constraint_force(*fw.joints[3], out)
    last = snaps[-1].joints
    assert (out[0], out[1]) == (inv_h * last["linear_impulse_x"][3], inv_h * last["linear_impulse_y"][3])
    assert L.f2d_joint_get_constraint_torque(*fw.joints[0]) == inv_h * last["spring_impulse"][0]
    L.f2d_joint_get_constraint_force(*fw.joints[4], out)
    assert (out[0], out[1]) == (0.0, 0.0)
    assert L.f2d_motor_joint_get_max_force(*fw.joints[1]) == 0.0
    assert L.f2d_motor_joint_get_correction_factor(*fw.joints[1]) == 1.0
    fw.close()


def test_destroy_joint_matches_reference(gpu):
    """f2d_destroy_joint (S/joint.c:702-819; fixture ops_destroy_joint: a mouse joint
    released and a new one made in its freed id, a motor and the filter joint destroyed,
    so the filtered boxes collide, a motor made in the next freed id, the first motor
    destroyed): after every one of 60 steps bodies, contacts, colours, islands and every
    joint slot (freed ones included) equal the reference's bit for bit; a handle to a
    destroyed joint is invalid once its id is reused."""
    fw, snaps = replay_joint_fixture("ops_destroy_joint", 5)
    import forge2d_amd as F
    L = F.lib()
    assert [L.f2d_joint_is_valid(*j) for j in fw.joints] == [0, 0, 0, 0, 1, 1, 1]
    assert [L.f2d_joint_get_type(*j) for j in fw.joints[4:]] == [3, 3, 2]
    fw.close()


def test_collide_connected_matches_reference(gpu):
    """f2d_joint_set_collide_connected (S/joint.c:979-1022; fixture ops_collide_connected:
    a box pulled into its carrier by a motor joint while they do not collide, then
    collideConnected turned on, so the box's shape enters the move buffer and the pair
    is created in the next step; a joint between distant boxes turned off and on):
    after every one of 60 steps the world equals the reference's bit for bit."""
    fw, snaps = replay_joint_fixture("ops_collide_connected", 2)
    import forge2d_amd as F
    L = F.lib()
    assert [L.f2d_joint_get_collide_connected(*j) for j in fw.joints] == [1, 1]
    fw.close()


def test_joint_between_touching_bodies_refused(gpu):
    """A joint with collideConnected off between bodies that share a contact would have
    the reference destroy that contact (S/joint.c:313-345); that edit is not restated,
    so the creation is refused with a null id. With collideConnected on, and for a mouse
    joint (whose creation destroys nothing), the joint is made and the world steps on."""
    spec, snaps, ops = load_ops_fixture("ops_motor_mouse", GOLDEN)
    fw = FlatWorld(spec)
    for _ in range(30):
        fw.step()
    import forge2d_amd as F
    L = F.lib()
    mine = fw.snapshot()
    touching = [(int(c["shape_a"]), int(c["shape_b"])) for c in mine.live_contacts()]
    assert (0, 6) in touching or (6, 0) in touching, touching  # box 6 (shape 6) rests on the ground
    ground, box = fw.bodies[0], fw.bodies[6]
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_revolute_joint(fw.w, *ground, *box, 9.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0, 0.0, 0,
                                0.0, 0.0, 1.0, 0, out)
    assert out[0] == 0
    L.f2d_create_revolute_joint(fw.w, *ground, *box, 9.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0, 0.0, 0,
                                0.0, 0.0, 1.0, 1, out)
    assert out[0] > 0
    L.f2d_create_mouse_joint(fw.w, *ground, *box, 9.0, 1.0, 5.0, 0.7, 100.0, 0, out)
    assert out[0] > 0
    fw.step()
    assert L.f2d_body_get_joint_count(*box) == 2
    fw.close()


def test_distance_joints_match_reference(gpu):
    """Distance joints (S/distance_joint.c) made through f2d_create_distance_joint and the
    f2d_distance_joint_* calls (fixture ops_distance: a rigid rope, a spring with a
    length range, a motor that drives two boxes into the upper and then, reversed, the
    lower end of their range with the spring branch on at zero stiffness, a spring with
    collideConnected; later a new rest length, the spring turned off (rigid branch) and a
    tightened range): after every one of 60 steps bodies, contacts, colours, islands and
    every joint's impulses equal the reference's bit for bit."""
    fw, snaps = replay_joint_fixture("ops_distance", 4)
    import forge2d_amd as F
    L = F.lib()
    assert [L.f2d_joint_get_type(*j) for j in fw.joints] == [0, 0, 0, 0]
    assert L.f2d_distance_joint_get_length(*fw.joints[0]) == 2.0
    assert (L.f2d_distance_joint_get_min_length(*fw.joints[3]), L.f2d_distance_joint_get_max_length(*fw.joints[3])) \
        == (np.float32(1.0), np.float32(1.1))
    inv_h = np.float32(4.0) * (np.float32(1.0) / np.float32(1.0 / 60.0))
    assert L.f2d_distance_joint_get_motor_force(*fw.joints[2]) == inv_h * snaps[-1].joints["motor_impulse"][2]
    fw.close()


def test_prismatic_joints_match_reference(gpu):
    """Prismatic joints (S/prismatic_joint.c) made through f2d_create_prismatic_joint and
    the f2d_prismatic_joint_* setters (fixture ops_prismatic: a vertical slider with
    limit and motor, an axis given unnormalized with a spring and a target, a joint
    between two dynamic boxes, one with collideConnected and limit, motor and spring
    together, one on a fixed-rotation body; later a new target, a reversed motor,
    moved limits and limit and spring turned off): after every one of 60 steps the
    bodies, contacts, colours, islands and every joint's impulses equal the
    reference's bit for bit."""
    spec, snaps, ops = load_ops_fixture("ops_prismatic", GOLDEN)
    fw = FlatWorld(spec)
    k = 0
    for ref in snaps[1:]:
        s = ref.step
        while k < len(ops) and int(ops[k]["step"]) == s:
            fw.apply(ops[k])
            k += 1
        print(f"[prismatic] step {s}", flush=True)
        fw.step()
        mine = fw.snapshot()
        for f in P.BODY_FIELDS + ["set_index", "local_index"]:
            assert np.array_equal(mine.bodies[f], ref.bodies[f]), (s, f)
        assert len(mine.joints) == len(ref.joints) == 5, s
        for f in JOINT_FIELDS:
            assert np.array_equal(mine.joints[f], ref.joints[f]), (s, f)
        assert contacts_by_id(mine) == contacts_by_id(ref), s
        rep = P.manifold_report(mine, ref, with_impulses=True)
        assert rep["point_count"] == rep["ids"] == rep["normals"] == rep["values"] == rep["impulses"] == 0, (s, rep)
        assert all(np.array_equal(a, c) for a, c in zip(mine.colors, ref.colors)), s
        assert P.island_differences(mine, ref) == [], s
    assert k == len(ops)
    # the accessors read back what the calls set
    import forge2d_amd as F
    L = F.lib()
    j1 = fw.joints[1]
    assert L.f2d_joint_get_type(*j1) == 4
    assert L.f2d_prismatic_joint_get_target_translation(*j1) == np.float32(-0.5)
    assert L.f2d_prismatic_joint_is_spring_enabled(*j1) == 1
    j0 = fw.joints[0]
    assert L.f2d_prismatic_joint_get_motor_speed(*j0) == np.float32(-1.5)
    assert L.f2d_prismatic_joint_get_lower_limit(*j0) == np.float32(-1.0)
    assert L.f2d_prismatic_joint_get_upper_limit(*j0) == np.float32(2.0)
    # the world's inv_h (subStepCount * (1 / timeStep) in float, S/world.c:760-771) times the
    # last motor impulse (S/prismatic_joint.c:138-144)
    inv_h = np.float32(4.0) * (np.float32(1.0) / np.float32(1.0 / 60.0))
    assert L.f2d_prismatic_joint_get_motor_force(*j0) == inv_h * snaps[-1].joints["motor_impulse"][0]
    fw.close()

