"""The flat per-world ABI on the host (no device step needed):

* a world built call by call through f2d_create_world / f2d_create_body /
  f2d_create_*_shape / f2d_create_revolute_joint equals the reference's own world
  after creation (step-0 snapshot of every fixture: bodies, shapes, fat boxes, trees,
  islands, joints), so a binding that builds its worlds through the flat ABI starts
  from the reference's state bit for bit;
* forge2d's own accessor tests replayed through the flat ABI with their tolerances:
  F/test/api/body_test.dart:90-150, 169-175, F/test/api/accessors_test.dart:26-78,
  F/test/api/shape_test.dart:94-183 (minus chains, sensors and user data, which have
  no counterpart here), plus the joint accessors of a revolute joint.
"""
import ctypes
import math

import numpy as np
import pytest

import forge2d_amd as F
from tests import parity_util as P
from tests.conftest import FIXTURE_NAMES, GOLDEN, load_fixture
from tests.flat_util import FlatWorld, ShapeDef, load_ops_fixture


@pytest.mark.parametrize("name", FIXTURE_NAMES + ["ops_bodies"])
def test_flat_build_matches_reference_step0(name):
    if name == "ops_bodies":
        spec, snaps, _ = load_ops_fixture(name, GOLDEN)
    else:
        spec, snaps, _ = load_fixture(name)
    assert snaps[0].step == 0
    fw = FlatWorld(spec)
    mine, ref = fw.snapshot(), snaps[0]
    for f in P.BODY_FIELDS + ["c0x", "c0y", "q0c", "q0s", "lcx", "lcy", "inv_mass", "inv_i", "min_ext", "max_ext",
                              "type", "set_index", "local_index", "flags"]:
        assert np.array_equal(mine.bodies[f], ref.bodies[f]), f
    for f in ("aabb", "fat", "proxy_key", "body"):
        assert np.array_equal(mine.shapes[f], ref.shapes[f]), f
    assert np.array_equal(mine.move, ref.move)
    assert P.tree_differences(mine, ref) == []
    assert P.island_differences(mine, ref) == []
    fw.close()


def vec(fn, *args, n=2):
    out = (ctypes.c_float * n)()
    fn(*args, out)
    return list(out)


def box_def(density=1.0):
    d = ShapeDef()
    d.friction, d.density = 0.6, density
    d.category_lo, d.mask_lo, d.mask_hi = 1, 0xFFFFFFFF, 0xFFFFFFFF
    d.invoke_contact_creation = d.update_body_mass = 1
    return d


def make_body(w, type_=2, x=0.0, y=0.0, angle=0.0):
    out = (ctypes.c_int32 * 2)()
    F.lib().f2d_create_body(w, type_, x, y, math.cos(angle), math.sin(angle), 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.05,
                            None, 1, 1, 0, 0, 1, 0, out)
    assert out[0] > 0
    return out[0], ctypes.c_uint32(out[1]).value


def add_box(b, h=0.5, density=1.0):
    out = (ctypes.c_int32 * 2)()
    F.lib().f2d_create_box_shape(b[0], b[1], ctypes.byref(box_def(density)), h, h, 0.0, 0.0, 1.0, 0.0, 0.0, out)
    return out[0], ctypes.c_uint32(out[1]).value


def new_world(gravity=(0.0, -10.0)):
    return F.lib().f2d_create_world(gravity[0], gravity[1], 1.0, 1.0, 30.0, 10.0, 3.0, 400.0, 1, 1)


def test_body_mass_and_mass_data():
    L = F.lib()
    w = new_world()
    b = make_body(w)
    add_box(b, 0.5, density=2.0)
    assert abs(L.f2d_body_get_mass(*b) - 2.0) < 1e-6 and L.f2d_body_get_rotational_inertia(*b) > 0  # body_test:90-96
    original = L.f2d_body_get_mass(*b)
    L.f2d_body_set_mass_data(*b, 42.0, 3.0, 0.1, 0.0)
    assert abs(L.f2d_body_get_mass(*b) - 42.0) < 1e-4                                              # :98-112
    assert abs(vec(L.f2d_body_get_local_center, *b)[0] - 0.1) < 1e-6
    L.f2d_body_apply_mass_from_shapes(*b)
    assert abs(L.f2d_body_get_mass(*b) - original) < 1e-6
    L.f2d_destroy_world(w)


def test_impulses_change_velocity_immediately():
    L = F.lib()
    w = new_world()
    b = make_body(w)
    add_box(b)
    L.f2d_body_apply_linear_impulse(*b, L.f2d_body_get_mass(*b), 0.0, 0.0, 0.0, 1)                # body_test:115-123
    assert abs(vec(L.f2d_body_get_linear_velocity, *b)[0] - 1.0) < 1e-5
    L.f2d_body_apply_angular_impulse(*b, L.f2d_body_get_rotational_inertia(*b), 1)
    assert abs(L.f2d_body_get_angular_velocity(*b) - 1.0) < 1e-5
    L.f2d_destroy_world(w)


def test_transform_and_point_conversions():
    L = F.lib()
    w = new_world()
    b = make_body(w, type_=0)
    L.f2d_body_set_transform(*b, 5.0, 6.0, math.cos(1.0), math.sin(1.0))                          # body_test:134-140
    p, q = vec(L.f2d_body_get_position, *b), vec(L.f2d_body_get_rotation, *b)
    assert abs(p[0] - 5) < 1e-6 and abs(p[1] - 6) < 1e-6 and abs(math.atan2(q[1], q[0]) - 1.0) < 1e-6
    c = make_body(w, type_=0, x=2.0, y=3.0, angle=0.7)                                              # :142-150
    wp = (ctypes.c_float * 2)()
    L.f2d_body_get_world_point(*c, 1.0, 1.0, wp)
    lp = vec(L.f2d_body_get_local_point, *c, wp[0], wp[1])
    assert abs(lp[0] - 1) < 1e-5 and abs(lp[1] - 1) < 1e-5
    L.f2d_destroy_world(w)


def test_shapes_list_all_attached_shapes():
    L = F.lib()
    w = new_world()
    b = make_body(w, type_=0)
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_circle_shape(*b, ctypes.byref(box_def()), 0.0, 0.0, 1.0, out)
    first = (out[0], ctypes.c_uint32(out[1]).value)
    second = add_box(b, 1.0)                                                                        # body_test:169-175
    ids = (ctypes.c_int32 * 8)()
    n = L.f2d_body_get_shapes(*b, ids, 4)
    assert n == L.f2d_body_get_shape_count(*b) == 2
    got = {(ids[2 * i], ctypes.c_uint32(ids[2 * i + 1]).value) for i in range(n)}
    assert got == {first, second}
    L.f2d_destroy_world(w)


def test_body_accessors_round_trip():
    L = F.lib()
    w = new_world()
    b = make_body(w, y=5.0)
    add_box(b)
    L.f2d_body_set_angular_velocity(*b, 1.25)                                                       # accessors_test:26-63
    L.f2d_body_set_linear_velocity(*b, 2.0, 3.0)
    L.f2d_body_set_gravity_scale(*b, 0.5)
    L.f2d_body_set_linear_damping(*b, 0.125)
    L.f2d_body_set_angular_damping(*b, 0.25)
    L.f2d_body_set_sleep_threshold(*b, 0.75)
    L.f2d_body_enable_sleep(*b, 0)
    L.f2d_body_set_bullet(*b, 1)
    L.f2d_body_set_name(*b, b"renamed")
    assert abs(L.f2d_body_get_angular_velocity(*b) - 1.25) < 1e-6
    assert abs(vec(L.f2d_body_get_linear_velocity, *b)[0] - 2) < 1e-6
    assert abs(L.f2d_body_get_gravity_scale(*b) - 0.5) < 1e-6
    assert abs(L.f2d_body_get_linear_damping(*b) - 0.125) < 1e-6
    assert abs(L.f2d_body_get_angular_damping(*b) - 0.25) < 1e-6
    assert abs(L.f2d_body_get_sleep_threshold(*b) - 0.75) < 1e-6
    assert L.f2d_body_is_sleep_enabled(*b) == 0 and L.f2d_body_is_bullet(*b) == 1
    L.f2d_body_set_fixed_rotation(*b, 1)
    assert L.f2d_body_is_fixed_rotation(*b) == 1 and L.f2d_body_get_angular_velocity(*b) == 0.0
    assert L.f2d_body_get_name(*b) == b"renamed"
    assert abs(vec(L.f2d_body_get_position, *b)[1] - 5) < 1e-6
    assert abs(vec(L.f2d_body_get_world_center, *b)[1] - 5) < 1e-6
    assert L.f2d_body_is_enabled(*b) == 1
    L.f2d_destroy_world(w)


def test_shape_accessors_round_trip():
    L = F.lib()
    w = new_world()
    b = make_body(w, y=5.0)
    s = add_box(b)
    L.f2d_shape_enable_pre_solve_events(*s, 1)                                                     # accessors_test:65-78
    assert L.f2d_shape_are_pre_solve_events_enabled(*s) == 1
    L.f2d_shape_set_density(*s, 4.0, 0)
    assert L.f2d_shape_get_density(*s) == 4.0
    L.f2d_body_apply_mass_from_shapes(*b)
    assert abs(L.f2d_body_get_mass(*b) - 4.0) < 1e-5
    L.f2d_shape_set_friction(*s, 0.25)                                                              # shape_test:94-114
    L.f2d_shape_set_restitution(*s, 0.75)
    L.f2d_shape_enable_contact_events(*s, 1)
    L.f2d_shape_enable_hit_events(*s, 1)
    assert abs(L.f2d_shape_get_friction(*s) - 0.25) < 1e-6
    assert abs(L.f2d_shape_get_restitution(*s) - 0.75) < 1e-6
    assert L.f2d_shape_are_contact_events_enabled(*s) == 1 and L.f2d_shape_are_hit_events_enabled(*s) == 1
    filt = (ctypes.c_int32 * 5)()
    L.f2d_shape_get_filter(*s, filt)
    assert filt[0] == 1 and ctypes.c_uint32(filt[2]).value == 0xFFFFFFFF
    body = (ctypes.c_int32 * 2)()
    L.f2d_shape_get_body(*s, body)
    assert (body[0], ctypes.c_uint32(body[1]).value) == b
    L.f2d_destroy_world(w)


def test_geometry_reads_back_every_shape_kind():
    L = F.lib()
    w = new_world()
    b = make_body(w, type_=0)
    d = ctypes.byref(box_def())
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_circle_shape(*b, d, 1.0, 2.0, 0.75, out)                                           # shape_test:116-170
    assert L.f2d_shape_get_type(out[0], out[1]) == 0
    assert np.allclose(vec(L.f2d_shape_get_circle, out[0], out[1], n=3), [1, 2, 0.75], atol=1e-6)
    L.f2d_create_capsule_shape(*b, d, 0.0, -1.0, 0.0, 1.0, 0.5, out)
    assert L.f2d_shape_get_type(out[0], out[1]) == 1
    assert np.allclose(vec(L.f2d_shape_get_capsule, out[0], out[1], n=5), [0, -1, 0, 1, 0.5], atol=1e-6)
    L.f2d_create_segment_shape(*b, d, -2.0, 0.0, 2.0, 1.0, out)
    assert L.f2d_shape_get_type(out[0], out[1]) == 2
    assert np.allclose(vec(L.f2d_shape_get_segment, out[0], out[1], n=4), [-2, 0, 2, 1], atol=1e-6)
    L.f2d_create_box_shape(*b, d, 1.5, 0.5, 0.0, 0.0, 1.0, 0.0, 0.0, out)
    poly = (ctypes.c_float * 17)()
    assert L.f2d_shape_get_type(out[0], out[1]) == 3
    assert L.f2d_shape_get_polygon(out[0], out[1], poly) == 4 and poly[0] == 0.0
    assert all(abs(abs(poly[1 + 2 * i]) - 1.5) < 1e-5 and abs(abs(poly[2 + 2 * i]) - 0.5) < 1e-5 for i in range(4))
    L.f2d_destroy_world(w)


def test_test_point_and_aabb():
    L = F.lib()
    w = new_world()
    b = make_body(w, type_=0)
    s = add_box(b, 1.0)                                                                             # shape_test:175-183
    assert L.f2d_shape_test_point(*s, 0.5, 0.5) == 1 and L.f2d_shape_test_point(*s, 2.0, 2.0) == 0
    a = vec(L.f2d_shape_get_aabb, *s, n=4)
    assert abs(a[0] + 1) < 0.1 and abs(a[3] - 1) < 0.1
    c = make_body(w, type_=0, x=5.0)
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_capsule_shape(*c, ctypes.byref(box_def()), 0.0, -1.0, 0.0, 1.0, 0.5, out)
    assert L.f2d_shape_test_point(out[0], out[1], 5.2, 1.3) == 1
    assert L.f2d_shape_test_point(out[0], out[1], 5.6, 0.0) == 0
    L.f2d_destroy_world(w)


def test_revolute_joint_accessors():
    L = F.lib()
    w = new_world()
    g = make_body(w, type_=0)
    b = make_body(w, y=-1.0)
    add_box(b, 0.2)
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_revolute_joint(w, *g, *b, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0, 0.0,
                                0.25, 1, out)
    j = (out[0], ctypes.c_uint32(out[1]).value)
    assert L.f2d_joint_is_valid(*j) == 1 and L.f2d_joint_get_type(*j) == 5
    ba, bb = (ctypes.c_int32 * 2)(), (ctypes.c_int32 * 2)()
    L.f2d_joint_get_body_a(*j, ba)
    L.f2d_joint_get_body_b(*j, bb)
    assert (ba[0], ctypes.c_uint32(ba[1]).value) == g and (bb[0], ctypes.c_uint32(bb[1]).value) == b
    assert vec(L.f2d_joint_get_local_anchor_b, *j) == [0.0, 1.0]
    assert L.f2d_joint_get_collide_connected(*j) == 1
    ids = (ctypes.c_int32 * 4)()
    assert L.f2d_body_get_joint_count(*b) == 1 and L.f2d_body_get_joints(*b, ids, 2) == 1
    assert (ids[0], ctypes.c_uint32(ids[1]).value) == j
    L.f2d_destroy_world(w)


def test_ops_fixture_exercises_every_call():
    spec, snaps, ops = load_ops_fixture("ops_bodies", GOLDEN)
    assert set(int(o) for o in ops["op"]) == {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 23}
    by = {s.step: s for s in snaps}
    # the body created asleep stays asleep until the call with wake (before step 20)
    assert by[19].bodies["set_index"][8] >= 3 and by[20].bodies["set_index"][8] == 2
    # the static post's proxy moved (before step 8)
    assert not np.array_equal(by[7].shapes["fat"][1], by[8].shapes["fat"][1])
    # the impulse beyond the speed limit is clipped to the world's 400 m/s (before step 21)
    assert np.hypot(by[21].bodies["vx"][5], by[21].bodies["vy"][5]) < 400.0


def test_prismatic_joint_accessors_on_the_host():
    """f2d_create_prismatic_joint and the f2d_prismatic_joint_* calls on the host
    (S/prismatic_joint.c:12-209): the step-1 calls of the ops_prismatic fixture made
    through the flat ABI read back as set (limits ordered by SetLimits, enable flags,
    target, motor and spring parameters), the joint type is b2_prismaticJoint (4), a
    call on a revolute joint is ignored, and the translation along the normalized
    axis matches the bodies' placement."""
    spec, snaps, ops = load_ops_fixture("ops_prismatic", GOLDEN)
    fw = FlatWorld(spec)
    for op in ops[ops["step"] == 1]:
        fw.apply(op)
    L = F.lib()
    assert len(fw.joints) == 5
    assert [L.f2d_joint_get_type(*j) for j in fw.joints] == [4] * 5
    j0, j1, j2, j3 = fw.joints[:4]
    assert (L.f2d_prismatic_joint_is_limit_enabled(*j0), L.f2d_prismatic_joint_is_motor_enabled(*j0),
            L.f2d_prismatic_joint_is_spring_enabled(*j0)) == (1, 1, 0)
    assert (L.f2d_prismatic_joint_get_lower_limit(*j0), L.f2d_prismatic_joint_get_upper_limit(*j0)) == (-1.0, 2.0)
    assert L.f2d_prismatic_joint_get_max_motor_force(*j0) == 50.0 and L.f2d_prismatic_joint_get_motor_speed(*j0) == 1.0
    assert L.f2d_prismatic_joint_get_spring_hertz(*j1) == 2.0
    assert L.f2d_prismatic_joint_get_spring_damping_ratio(*j1) == 0.5
    assert L.f2d_prismatic_joint_get_target_translation(*j1) == 0.5
    assert L.f2d_joint_get_collide_connected(*j3) == 1 and L.f2d_joint_get_collide_connected(*j2) == 0
    L.f2d_prismatic_joint_set_limits(*j2, 3.0, -1.0)  # SetLimits orders them
    assert (L.f2d_prismatic_joint_get_lower_limit(*j2), L.f2d_prismatic_joint_get_upper_limit(*j2)) == (-1.0, 3.0)
    # joint 0: ground (0, -1) anchor (-6, 0) -> (-6, -1); body 1 at (-6, 3) anchor (0, -1) -> (-6, 2): 3 along +y
    assert math.isclose(L.f2d_prismatic_joint_get_translation(*j0), 3.0, abs_tol=1e-6)
    # joint 1's axis (1, 1) is normalized: ground anchor (-2, 1) -> (-2, 0), body 2's centre (-2, 2): 2 / sqrt(2)
    assert math.isclose(L.f2d_prismatic_joint_get_translation(*j1), math.sqrt(2.0), rel_tol=1e-6)
    fw.close()


def test_prismatic_fixture_exercises_every_branch():
    """Over the 60 reference steps of ops_prismatic the joints hold lower and upper limit
    impulses, motor impulses of both signs, spring impulses, an angular block impulse,
    and the fixed-rotation joint (k22 = 0 replaced by 1) a perpendicular impulse."""
    spec, snaps, ops = load_ops_fixture("ops_prismatic", GOLDEN)
    J = np.stack([s.joints for s in snaps[1:]])
    assert (J["reserved"] == 1.0).all()
    assert (J["lower_impulse"] > 0).any() and (J["upper_impulse"] > 0).any()
    assert (J["motor_impulse"] > 0).any() and (J["motor_impulse"] < 0).any()
    assert (J["spring_impulse"] != 0).any()
    assert (J["linear_impulse_y"] != 0).any()
    assert (J["linear_impulse_x"][:, 4] != 0).any()


def test_weld_wheel_fixture_exercises_every_branch():
    """Over the 60 reference steps of ops_weld_wheel: weld linear and angular impulses
    (rigid and soft), and on the rear wheel motor impulses of both signs, lower and
    upper limit impulses and spring impulses; the front wheel's perpendicular impulse."""
    spec, snaps, ops = load_ops_fixture("ops_weld_wheel", GOLDEN)
    J = np.stack([s.joints for s in snaps[1:]])
    assert J["reserved"][0].tolist() == [2.0, 2.0, 3.0, 3.0, 2.0]
    assert (J["linear_impulse_y"][:, 0] != 0).any() and (J["spring_impulse"][:, 1] != 0).any()
    assert (J["motor_impulse"][:, 2] > 0).any() and (J["motor_impulse"][:, 2] < 0).any()
    assert (J["lower_impulse"][:, 2] > 0).any() and (J["upper_impulse"][:, 2] > 0).any()
    assert (J["spring_impulse"][:, 2] != 0).any() and (J["linear_impulse_x"][:, 3] != 0).any()


def test_distance_fixture_exercises_every_branch():
    """Over the 60 reference steps of ops_distance: the rigid branch's impulse (rope, and
    the spring joint after its spring is turned off), the spring branch, motor impulses
    of both signs and lower and upper range impulses."""
    spec, snaps, ops = load_ops_fixture("ops_distance", GOLDEN)
    J = np.stack([s.joints for s in snaps[1:]])
    assert J["reserved"][0].tolist() == [4.0] * 4
    assert (J["linear_impulse_x"][:, 0] != 0).any() and (J["linear_impulse_x"][:, 1] != 0).any()
    assert (J["motor_impulse"][:, 2] > 0).any() and (J["motor_impulse"][:, 2] < 0).any()
    assert (J["lower_impulse"][:, 2] > 0).any() and (J["upper_impulse"][:, 2] > 0).any()


def test_revolute_fixture_exercises_every_branch():
    """Over the 40 reference steps of ops_revolute: motor impulses of both signs on
    joint 0, lower and upper limit impulses on every joint, and spring impulses of both
    signs on joints 1 and 2."""
    spec, snaps, ops = load_ops_fixture("ops_revolute", GOLDEN)
    J = np.stack([s.joints for s in snaps[1:]])
    assert (J["reserved"] == 0.0).all()
    assert (J["motor_impulse"][:, 0] > 0).any() and (J["motor_impulse"][:, 0] < 0).any()
    assert ((J["lower_impulse"] > 0).any(axis=0)).all() and ((J["upper_impulse"] > 0).any(axis=0)).all()
    for k in (1, 2):
        assert (J["spring_impulse"][:, k] > 0).any() and (J["spring_impulse"][:, k] < 0).any()


def test_revolute_joint_calls():
    """The revolute parameter calls of f2d_shim.c:1097-1136 store what they are given
    (limits ordered as S/revolute_joint.c:113-127 orders them) and read it back; the
    joint angle of a fresh chain link is 0."""
    spec, snaps, ops = load_ops_fixture("ops_revolute", GOLDEN)
    fw = FlatWorld(spec)
    L = F.lib()
    j = fw.joints[1]
    assert L.f2d_revolute_joint_is_limit_enabled(*j) == 1 and L.f2d_revolute_joint_is_motor_enabled(*j) == 0
    assert (L.f2d_revolute_joint_get_lower_limit(*j), L.f2d_revolute_joint_get_upper_limit(*j)) == (-0.5, 0.5)
    L.f2d_revolute_joint_set_limits(*j, 0.4, -0.2)
    assert (L.f2d_revolute_joint_get_lower_limit(*j), L.f2d_revolute_joint_get_upper_limit(*j)) \
        == (np.float32(-0.2), np.float32(0.4))
    L.f2d_revolute_joint_enable_motor(*j, 1)
    L.f2d_revolute_joint_set_motor_speed(*j, 1.5)
    L.f2d_revolute_joint_set_max_motor_torque(*j, 7.0)
    L.f2d_revolute_joint_enable_spring(*j, 1)
    L.f2d_revolute_joint_set_spring_hertz(*j, 3.0)
    L.f2d_revolute_joint_set_spring_damping_ratio(*j, 0.25)
    L.f2d_revolute_joint_set_target_angle(*j, 0.125)
    assert L.f2d_revolute_joint_is_motor_enabled(*j) == 1 and L.f2d_revolute_joint_is_spring_enabled(*j) == 1
    assert (L.f2d_revolute_joint_get_motor_speed(*j), L.f2d_revolute_joint_get_max_motor_torque(*j)) == (1.5, 7.0)
    assert (L.f2d_revolute_joint_get_spring_hertz(*j), L.f2d_revolute_joint_get_spring_damping_ratio(*j)) \
        == (3.0, 0.25)
    assert L.f2d_revolute_joint_get_target_angle(*j) == 0.125
    assert L.f2d_revolute_joint_get_motor_torque(*j) == 0.0
    assert L.f2d_revolute_joint_get_angle(*j) == 0.0
    fw.close()


def test_motor_mouse_fixture_exercises_every_branch():
    """Over the 60 reference steps of ops_motor_mouse: the ground motor's linear impulse
    held on its force limit (|impulse| = h x max force), the carried motor's angular
    impulse on both torque limits and its linear impulse 0 once its force is set below 0
    (clamped to 0), the weak mouse joint on its force limit, and the filter joint with
    no impulse."""
    spec, snaps, ops = load_ops_fixture("ops_motor_mouse", GOLDEN)
    J = np.stack([s.joints for s in snaps[1:]])
    assert J["reserved"][0].tolist() == [5.0, 5.0, 6.0, 6.0, 7.0]
    h = np.float32(np.float32(1.0 / 60.0) / np.float32(4.0))
    mag0 = np.hypot(J["linear_impulse_x"][:, 0], J["linear_impulse_y"][:, 0])
    assert np.allclose(mag0[5:], np.float32(30.0) * h, rtol=1e-6)
    ang1 = J["spring_impulse"][:, 1]
    assert np.isclose(ang1, h).any() and np.isclose(ang1, -h).any()
    assert (J["linear_impulse_x"][31:, 1] == 0).all() and (J["linear_impulse_y"][31:, 1] == 0).all()
    mag3 = np.hypot(J["linear_impulse_x"][:, 3], J["linear_impulse_y"][:, 3])
    assert np.isclose(mag3[5:20], np.float32(3.0) * h, rtol=1e-6).all()
    assert (J["linear_impulse_x"][:, 4] == 0).all() and (J["spring_impulse"][:, 4] == 0).all()
    assert J["ref_angle"][-1].tolist()[:2] == [np.float32(-0.5), np.float32(0.4)]


def test_motor_mouse_joint_calls():
    """The motor and mouse calls of f2d_shim.c:1006-1050: values read back as stored,
    the motor's maximum force and torque clamped at 0 and its correction factor to
    [0, 1] (S/motor_joint.c:36-71); the mouse joint's local anchors are the target in
    each body's frame (S/joint.c:467-468); calls on the other type are ignored."""
    spec, snaps, ops = load_ops_fixture("ops_motor_mouse", GOLDEN)
    fw = FlatWorld(spec)
    for op in ops[ops["step"] == 1]:
        fw.apply(op)
    L = F.lib()
    motor, mouse = fw.joints[0], fw.joints[2]
    assert [L.f2d_joint_get_type(*j) for j in fw.joints] == [2, 2, 3, 3, 1]
    out = (ctypes.c_float * 2)()
    L.f2d_motor_joint_get_linear_offset(*motor, out)
    assert (out[0], out[1]) == (0.5, 2.0)
    assert L.f2d_motor_joint_get_angular_offset(*motor) == np.float32(0.8)
    L.f2d_motor_joint_set_max_force(*motor, -3.0)
    L.f2d_motor_joint_set_max_torque(*motor, -1.0)
    L.f2d_motor_joint_set_correction_factor(*motor, -0.5)
    assert (L.f2d_motor_joint_get_max_force(*motor), L.f2d_motor_joint_get_max_torque(*motor),
            L.f2d_motor_joint_get_correction_factor(*motor)) == (0.0, 0.0, 0.0)
    L.f2d_motor_joint_set_correction_factor(*motor, 2.0)
    assert L.f2d_motor_joint_get_correction_factor(*motor) == 1.0
    L.f2d_mouse_joint_get_target(*mouse, out)
    assert (out[0], out[1]) == (3.0, 0.5)
    L.f2d_mouse_joint_set_target(*mouse, -1.0, 2.5)
    L.f2d_mouse_joint_get_target(*mouse, out)
    assert (out[0], out[1]) == (-1.0, 2.5)
    assert (L.f2d_mouse_joint_get_spring_hertz(*mouse), L.f2d_mouse_joint_get_spring_damping_ratio(*mouse),
            L.f2d_mouse_joint_get_max_force(*mouse)) == (5.0, np.float32(0.7), 500.0)
    # the target (3, 0.5) in the frames of the ground (0, -1) and of body 4 at (3, 0.5)
    L.f2d_joint_get_local_anchor_a(*mouse, out)
    assert (out[0], out[1]) == (3.0, 1.5)
    L.f2d_joint_get_local_anchor_b(*mouse, out)
    assert (out[0], out[1]) == (0.0, 0.0)
    L.f2d_mouse_joint_set_max_force(*motor, 9.0)
    assert L.f2d_mouse_joint_get_max_force(*motor) == 0.0 and L.f2d_motor_joint_get_max_force(*mouse) == 0.0
    fw.close()


def test_destroy_joint_on_the_host():
    """f2d_destroy_joint before any step: the joint's slot is freed (id -1 in the
    snapshot), the colour's last joint takes its local index (S/constraint_graph.c:
    292-322), its bodies' joint lists drop it, the next joint created takes the freed id
    with a new generation (the old handle is invalid), as in the reference's fixture
    ops_destroy_joint (step 21: the joints of colour 0 at local indices 0, 1, -, 3, 2)."""
    spec, snaps, ops = load_ops_fixture("ops_destroy_joint", GOLDEN)
    fw = FlatWorld(spec)
    for op in ops[ops["step"] == 1]:
        fw.apply(op)
    L = F.lib()
    mouse = fw.joints[2]
    L.f2d_destroy_joint(*mouse)
    assert L.f2d_joint_is_valid(*mouse) == 0
    assert L.f2d_body_get_joint_count(*fw.bodies[4]) == 0
    snap = fw.snapshot()
    ref = snaps[21].joints
    assert snap.joints["id"].tolist() == ref["id"].tolist() == [0, 1, -1, 3, 4]
    assert snap.joints["local_index"].tolist() == ref["local_index"].tolist()
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_mouse_joint(fw.w, *fw.bodies[0], *fw.bodies[4], 2.0, 4.0, 5.0, 0.7, 500.0, 0, out)
    assert out[0] == 3 and (out[1] >> 16) == 2
    assert L.f2d_joint_is_valid(*mouse) == 0 and L.f2d_joint_is_valid(out[0], out[1]) == 1
    fw.close()


def test_collide_connected_fixture_creates_the_pair():
    """In ops_collide_connected the carrier and the box its motor joint pulls in (shapes
    2 and 3) share no contact until collideConnected is turned on before step 20, and
    share one from then on."""
    spec, snaps, ops = load_ops_fixture("ops_collide_connected", GOLDEN)
    def touching(st):
        pairs = {(int(c["shape_a"]), int(c["shape_b"])) for c in snaps[st].live_contacts()}
        return (2, 3) in pairs or (3, 2) in pairs
    assert not any(touching(st) for st in range(1, 20))
    assert all(touching(st) for st in range(20, 61))


def test_joint_calls_on_other_types_are_ignored():
    """A joint parameter call of the wrong joint type leaves the joint alone and its
    getter returns 0, as the reference's type check would refuse it."""
    spec, snaps, ops = load_ops_fixture("ops_distance", GOLDEN)
    fw = FlatWorld(spec)
    for op in ops[ops["step"] == 1]:
        fw.apply(op)
    L = F.lib()
    j = fw.joints[0]
    assert L.f2d_prismatic_joint_get_lower_limit(*j) == 0.0
    L.f2d_wheel_joint_set_motor_speed(*j, 3.0)
    L.f2d_prismatic_joint_set_limits(*j, -1.0, 1.0)
    assert L.f2d_distance_joint_get_min_length(*j) == np.float32(0.005)
    assert L.f2d_distance_joint_get_length(*j) == 3.0
    assert L.f2d_distance_joint_get_motor_speed(*j) == 0.0
    fw.close()
