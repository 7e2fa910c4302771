"""The flat per-world ABI (f2d_*, the reference shim's signatures) replaying the
reference's own behavioural tests with their tolerances:

* F/test/ffi/smoke_test.dart:13-46 and F/test/api/world_test.dart:56-74 — a 1x1 box
  dropped from y=10 rests on a 50x1 ground box at (0, 0.5) +- 0.01 after 200 steps
  (here with sleep off; tests/test_gpu_sleep.py replays it with the sleep assertion);
* F/test/api/body_test.dart:90-96 — a 1x1 box of density 2 has mass 2 +- 1e-6;
* F/test/api/body_test.dart:115-132 — an impulse of m*1 gives v = 1 +- 1e-5, and a
  force of m*60 for one 1/60 s step gives v = 1 +- 0.01 (gravity off);
* F/test/api/joints_test.dart:243-262 — a revolute joint limit of +-pi/8 holds
  within 0.02 after 240 steps.
"""
import ctypes
import math

import pytest

import forge2d_amd as F

pytestmark = pytest.mark.gpu

PI = ctypes.POINTER(ctypes.c_int32)
PF = ctypes.POINTER(ctypes.c_float)


class ShapeDef(ctypes.Structure):
    _fields_ = [(n, t) for n, t in [
        ("friction", ctypes.c_float), ("restitution", ctypes.c_float), ("rolling_resistance", ctypes.c_float),
        ("tangent_speed", ctypes.c_float), ("user_material_id", ctypes.c_int32), ("custom_color", ctypes.c_uint32),
        ("density", ctypes.c_float), ("category_lo", ctypes.c_uint32), ("category_hi", ctypes.c_uint32),
        ("mask_lo", ctypes.c_uint32), ("mask_hi", ctypes.c_uint32), ("group_index", ctypes.c_int32),
        ("is_sensor", ctypes.c_int32), ("enable_sensor_events", ctypes.c_int32),
        ("enable_contact_events", ctypes.c_int32), ("enable_hit_events", ctypes.c_int32),
        ("enable_pre_solve_events", ctypes.c_int32), ("invoke_contact_creation", ctypes.c_int32),
        ("update_body_mass", ctypes.c_int32)]]


def shape_def(density=1.0, friction=0.6):
    d = ShapeDef()
    d.friction = friction
    d.density = density
    d.category_lo = 1
    d.mask_lo = d.mask_hi = 0xFFFFFFFF
    d.invoke_contact_creation = 1
    d.update_body_mass = 1
    return d


def world(gravity=(0.0, -10.0)):
    L = F.lib()
    w = L.f2d_create_world(gravity[0], gravity[1], 1.0, 1.0, 30.0, 10.0, 3.0, 400.0, 0, 1)
    assert L.f2d_world_is_valid(w)
    return w


def body(w, type_, x, y):
    out = (ctypes.c_int32 * 2)()
    F.lib().f2d_create_body(w, type_, x, y, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.05, None, 1, 1, 0, 0, 1, 0, out)
    assert out[0] > 0
    return out[0], ctypes.c_uint32(out[1]).value


def box(bid, hw, hh, density=1.0):
    out = (ctypes.c_int32 * 2)()
    d = shape_def(density)
    F.lib().f2d_create_box_shape(bid[0], bid[1], ctypes.byref(d), hw, hh, 0.0, 0.0, 1.0, 0.0, 0.0, out)
    assert out[0] > 0


def pos(bid):
    out = (ctypes.c_float * 2)()
    F.lib().f2d_body_get_position(bid[0], bid[1], out)
    return out[0], out[1]


def vel(bid):
    out = (ctypes.c_float * 2)()
    F.lib().f2d_body_get_linear_velocity(bid[0], bid[1], out)
    return out[0], out[1]


def test_falling_box_rests_on_ground(gpu):
    L = F.lib()
    w = world()
    ground = body(w, 0, 0.0, -1.0)
    box(ground, 50.0, 1.0)
    b = body(w, 2, 0.0, 10.0)
    box(b, 0.5, 0.5)
    for _ in range(200):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    x, y = pos(b)
    assert abs(x) < 0.01 and abs(y - 0.5) < 0.01, (x, y)
    L.f2d_destroy_world(w)
    assert not L.f2d_world_is_valid(w)


def test_box_mass(gpu):
    L = F.lib()
    w = world()
    b = body(w, 2, 0.0, 0.0)
    box(b, 0.5, 0.5, density=2.0)
    assert abs(L.f2d_body_get_mass(b[0], b[1]) - 2.0) < 1e-6
    L.f2d_destroy_world(w)


def test_impulse_and_force_change_velocity(gpu):
    L = F.lib()
    w = world(gravity=(0.0, 0.0))
    b = body(w, 2, 0.0, 0.0)
    box(b, 0.5, 0.5)
    m = L.f2d_body_get_mass(b[0], b[1])
    L.f2d_body_apply_linear_impulse_to_center(b[0], b[1], m * 1.0, 0.0, 1)
    assert abs(vel(b)[0] - 1.0) < 1e-5
    c = body(w, 2, 5.0, 0.0)
    box(c, 0.5, 0.5)
    L.f2d_body_apply_force_to_center(c[0], c[1], m * 60.0, 0.0, 1)
    L.f2d_world_step(w, 1.0 / 60.0, 4)
    assert abs(vel(c)[0] - 1.0) < 0.01
    L.f2d_destroy_world(w)


def test_revolute_limit_holds(gpu):
    """F/test/api/joints_test.dart:12-20, 243-262: ground box 50x1 at (0,-1), a unit box
    at (2,5) pinned at ground-local (0,6) / body-local (-2,0) with limits +-pi/8."""
    L = F.lib()
    w = world()
    ground = body(w, 0, 0.0, -1.0)
    box(ground, 50.0, 1.0)
    b = body(w, 2, 2.0, 5.0)
    box(b, 0.5, 0.5)
    out = (ctypes.c_int32 * 2)()
    lim = math.pi / 8
    L.f2d_create_revolute_joint(w, ground[0], ground[1], b[0], b[1], 0.0, 6.0, -2.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 1,
                                -lim, lim, 0, 0.0, 0.0, 0.25, 0, out)
    assert out[0] > 0
    for _ in range(240):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    q = (ctypes.c_float * 2)()
    L.f2d_body_get_rotation(b[0], b[1], q)
    angle = math.atan2(q[1], q[0])
    assert -lim - 0.02 <= angle <= lim + 0.02, angle
    L.f2d_destroy_world(w)


def test_revolute_motor_spins_body(gpu):
    """F/test/api/joints_test.dart:264-282: motor speed 1 reached within 0.05."""
    L = F.lib()
    w = world()
    ground = body(w, 0, 0.0, -1.0)
    box(ground, 50.0, 1.0)
    b = body(w, 2, 0.0, 5.0)
    box(b, 0.5, 0.5)
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_revolute_joint(w, ground[0], ground[1], b[0], b[1], 0.0, 6.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 0,
                                0.0, 0.0, 1, 1000.0, 1.0, 0.25, 0, out)
    for _ in range(60):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    assert abs(L.f2d_body_get_angular_velocity(b[0], b[1]) - 1.0) < 0.05
    L.f2d_destroy_world(w)
