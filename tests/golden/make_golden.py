"""Regenerate the golden fixtures from the reference Box2D (test infrastructure).

Runs only in the build container, where /root/reference exists: it needs the
reference build of ``oracle/Makefile`` (``make -C oracle ref``). For each
fixture scene it writes the scene spec and the reference's own world snapshots
after the listed steps, compressed, to ``tests/golden/<name>.npz``:

    spec   uint8   the scene spec (forge2d_amd/scenes.py layout)
    snaps  uint8   concatenated snapshots (forge2d_amd/snapshot.py layout)
    steps  int32   the step index of each snapshot
    geoms  uint8   per shape, the reference's built geometry (148 B: int32 type,
                   count; float radius, centroid, 8 vertices, 8 normals)
    events uint8   ("events" and "sleep") per step: int32 [step, begin, end, hit,
                   move] counts, then float64 records in the flat shim's event layouts

Usage: python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from forge2d_amd import scenes  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def overflow_world():
    """A wide dynamic plank under many small circles: every circle contact shares
    the plank's body, so after ten colours the rest go to the overflow colour
    (S/constraint_graph.c:68-135)."""
    w = scenes.World(scenes.world_header(sleep=False, continuous=False))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 30.0, 1.0)
    plank = w.add_body(scenes.DYNAMIC, 0.0, 0.5)
    w.add_box(plank, 12.0, 0.5, density=2.0)
    for i in range(36):
        b = w.add_body(scenes.DYNAMIC, -11.0 + 0.62 * i, 1.31 + 0.02 * (i % 3))
        w.add_circle(b, 0.3, friction=0.4, restitution=0.0, rolling=0.05 if i % 4 == 0 else 0.0)
    return w


def bounce_world():
    """Circles, capsules and boxes with restitution dropped on a ground box, so the
    restitution stage (S/contact_solver.c:1969-2076) applies impulses."""
    rng = scenes.SplitMix64(77)
    w = scenes.World(scenes.world_header(sleep=False, continuous=False))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 20.0, 1.0, restitution=0.2)
    for i in range(24):
        x = -9.0 + 0.8 * i
        b = w.add_body(scenes.DYNAMIC, x, 2.0 + rng.uniform(0.0, 2.0), angle=rng.uniform(-1.0, 1.0))
        k = i % 3
        e = rng.uniform(0.3, 0.8)
        if k == 0:
            w.add_circle(b, 0.3, restitution=e)
        elif k == 1:
            w.add_capsule(b, (-0.2, 0.0), (0.2, 0.0), 0.15, restitution=e)
        else:
            w.add_box(b, 0.25, 0.25, restitution=e)
    return w


def ccd_world():
    """Fast bodies of every shape type fired at thin static segments and a static
    box (b2SolveContinuous, S/solver.c:390-541): without the time-of-impact sweep
    they would pass through in one step. Two bullets cross a dynamic stack, which
    only bullets sweep against."""
    rng = scenes.SplitMix64(4242)
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, 0.0)
    w.add_segment(g, (-20.0, 0.0), (20.0, 0.0))
    w.add_segment(g, (-20.0, 0.0), (-20.0, 30.0))
    w.add_segment(g, (20.0, 0.0), (20.0, 30.0))
    w.add_box(g, 2.0, 0.1, cx=8.0, cy=6.0, angle=0.3)
    for i in range(30):
        x = -16.0 + 1.1 * i
        y = 2.0 + rng.uniform(0.0, 6.0)
        vx = rng.uniform(-20.0, 20.0) if i % 5 == 0 else 0.0
        b = w.add_body(scenes.DYNAMIC, x, y, angle=rng.uniform(-3.0, 3.0), vx=vx, vy=-rng.uniform(40.0, 120.0),
                       w=rng.uniform(-20.0, 20.0))
        k = i % 4
        if k == 0:
            w.add_circle(b, rng.uniform(0.1, 0.3))
        elif k == 1:
            w.add_capsule(b, (-0.3, 0.0), (0.3, 0.0), rng.uniform(0.08, 0.15))
        elif k == 2:
            w.add_box(b, rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.3))
        else:
            n = 3 + i % 6
            pts = []
            for j in range(n):
                a = 2.0 * 3.141592653589793 * j / n + rng.uniform(-0.2, 0.2)
                r = rng.uniform(0.15, 0.35)
                pts.append((r * scenes.math.cos(a), r * scenes.math.sin(a)))
            w.add_polygon(b, pts)
    # a dynamic stack and two bullets crossing it
    for r in range(6):
        b = w.add_body(scenes.DYNAMIC, -4.0, 0.5 + 1.0 * r)
        w.add_box(b, 0.5, 0.5)
    for j, (y, vx) in enumerate(((1.5, 150.0), (3.5, -180.0))):
        b = w.add_body(scenes.DYNAMIC, -12.0 if vx > 0 else 4.0, y, vx=vx, flags=scenes.BF_BULLET | scenes.BF_AWAKE)
        w.add_circle(b, 0.1, density=4.0)
    return w


def events_world():
    """Bouncing circles, capsules and boxes with contact and hit events enabled on
    every shape (S/world.c:630-676 begin/end events, S/solver.c:1758-1817 hits)."""
    rng = scenes.SplitMix64(99)
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    ev = scenes.SF_CONTACT_EVENTS | scenes.SF_HIT_EVENTS | scenes.SF_INVOKE_CONTACT | scenes.SF_UPDATE_MASS
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 20.0, 1.0, restitution=0.5, flags=ev)
    for i in range(20):
        x = -9.5 + 1.0 * i
        b = w.add_body(scenes.DYNAMIC, x, 1.0 + rng.uniform(0.0, 3.0), angle=rng.uniform(-1.0, 1.0),
                       vx=rng.uniform(-2.0, 2.0))
        k = i % 3
        e = rng.uniform(0.4, 0.9)
        if k == 0:
            w.add_circle(b, 0.3, restitution=e, flags=ev)
        elif k == 1:
            w.add_capsule(b, (-0.25, 0.0), (0.25, 0.0), 0.15, restitution=e, flags=ev)
        else:
            w.add_box(b, 0.3, 0.3, restitution=e, flags=ev)
    return w


def sleep_world():
    """Islands and sleep (S/island.c, S/solver_set.c, S/solver.c:1995-2049) with
    sleep enabled: two box pyramids settle and fall asleep as islands; a ball
    dropped from 30 m lands on one and wakes it (b2LinkContact wakes the sleeping
    solver set); a box slides into another and bounces off, so their merged island
    loses a contact and is split before it sleeps (S/island.c:602-830); a hanging
    revolute chain sleeps with its joints and is woken by a second ball."""
    w = scenes.World(scenes.world_header(sleep=True, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 40.0, 1.0, friction=0.6)
    for cx, rows in ((-10.0, 5), (10.0, 4)):
        for r in range(rows):
            n = rows - r
            for i in range(n):
                b = w.add_body(scenes.DYNAMIC, cx + (i - (n - 1) / 2.0) * 1.0, 0.5 + r * 1.0)
                w.add_box(b, 0.5, 0.5, friction=0.6)
    a = w.add_body(scenes.DYNAMIC, 0.0, 0.5)
    w.add_box(a, 0.5, 0.5, friction=0.6, restitution=0.6)
    s = w.add_body(scenes.DYNAMIC, -2.5, 0.5, vx=5.0)
    w.add_box(s, 0.5, 0.5, friction=0.6, restitution=0.6)
    anchor = w.add_body(scenes.STATIC, 20.0, 10.0)
    prev, prev_anchor = anchor, (0.0, 0.0)
    for k in range(4):
        b = w.add_body(scenes.DYNAMIC, 20.0, 10.0 - 0.2 - 0.4 * k)
        w.add_box(b, 0.05, 0.2, friction=0.6)
        w.add_revolute(prev, b, prev_anchor, (0.0, 0.2), enable_limit=True, lower=-0.5, upper=0.5)
        prev, prev_anchor = b, (0.0, -0.2)
    ball = w.add_body(scenes.DYNAMIC, -10.2, 30.0)
    w.add_circle(ball, 0.5, density=2.0, friction=0.6)
    ball2 = w.add_body(scenes.DYNAMIC, 20.3, 40.0)
    w.add_circle(ball2, 0.3, density=1.0, friction=0.6)
    return w


def features_world():
    """Body, shape and joint options that the other fixtures leave at their
    defaults: kinematic bodies (a moving platform and a spinner), linear and
    angular damping, gravity scales of 0.5, 2 and -0.3, fixed rotation, fast
    rotation, a conveyor (tangent speed, S/contact_solver.c), collision filter
    groups and category masks (S/shape.c b2ShouldShapesCollide), rounded boxes and
    polygons, and revolute joints with a motor, a spring, a motor within limits and
    collideConnected (S/revolute_joint.c:202-465)."""
    rng = scenes.SplitMix64(5151)
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 40.0, 1.0, friction=0.7)
    ceiling = w.add_body(scenes.STATIC, 0.0, 0.0)
    w.add_segment(ceiling, (20.0, 9.0), (30.0, 9.0))
    w.add_segment(ceiling, (-30.0, 3.0), (-22.0, 1.0))  # ramp
    # conveyor belt: a static box whose surface moves at 2 m/s
    conv = w.add_body(scenes.STATIC, -14.0, 2.0)
    w.add_box(conv, 4.0, 0.25, friction=0.8, tangent_speed=2.0)
    for i in range(4):
        b = w.add_body(scenes.DYNAMIC, -16.5 + 1.2 * i, 2.9 + 0.3 * i)
        w.add_box(b, 0.4, 0.4, friction=0.8, radius=0.05 * (i % 2))
    # kinematic platform carrying boxes, and a kinematic spinner
    plat = w.add_body(scenes.KINEMATIC, 0.0, 3.0, vx=0.5)
    w.add_box(plat, 3.0, 0.2)
    for i in range(3):
        b = w.add_body(scenes.DYNAMIC, -1.5 + 1.5 * i, 3.75)
        w.add_box(b, 0.5, 0.5, friction=0.6)
    spin = w.add_body(scenes.KINEMATIC, 8.0, 1.5, w=1.5)
    w.add_capsule(spin, (-1.5, 0.0), (1.5, 0.0), 0.2)
    for i in range(4):
        b = w.add_body(scenes.DYNAMIC, 6.8 + 0.8 * i, 3.0 + 0.6 * i)
        w.add_circle(b, 0.25)
    # damping and gravity scale
    for i, (ld, ad) in enumerate(((0.8, 1.5), (0.2, 0.0), (0.0, 4.0))):
        b = w.add_body(scenes.DYNAMIC, 12.0 + 1.5 * i, 2.0, vx=4.0 - 3.0 * i, w=5.0, lin_damp=ld, ang_damp=ad)
        w.add_circle(b, 0.3)
    for i, gs in enumerate((0.5, 2.0, -0.3)):
        b = w.add_body(scenes.DYNAMIC, 22.0 + 2.0 * i, 5.0, gravity_scale=gs, angle=0.3 * i)
        w.add_box(b, 0.4, 0.3)
    # fixed rotation on the ramp; fast rotation
    b = w.add_body(scenes.DYNAMIC, -27.0, 4.0, flags=scenes.BF_FIXED_ROT | scenes.BF_AWAKE | scenes.BF_ENABLE_SLEEP)
    w.add_box(b, 0.5, 0.25)
    b = w.add_body(scenes.DYNAMIC, -24.5, 4.0, angle=0.5, flags=scenes.BF_FIXED_ROT | scenes.BF_AWAKE)
    w.add_capsule(b, (-0.3, 0.0), (0.3, 0.0), 0.2)
    b = w.add_body(scenes.DYNAMIC, 17.0, 1.0, w=60.0, flags=scenes.BF_FAST_ROT | scenes.BF_AWAKE)
    w.add_box(b, 0.6, 0.1)
    # filter groups: a negative group never collides, a positive group always does
    for i in range(2):
        b = w.add_body(scenes.DYNAMIC, -6.0 + 0.3 * i, 1.0 + 0.9 * i)
        w.add_box(b, 0.5, 0.5, group=-3)
    for i in range(2):
        b = w.add_body(scenes.DYNAMIC, -8.5 + 0.2 * i, 1.0 + 1.0 * i)
        w.add_box(b, 0.5, 0.5, group=4, category=(2, 0), mask=(1, 0))
    # category masks: category 4 ignores category 2; a circle that masks out the ground
    b = w.add_body(scenes.DYNAMIC, -8.4, 4.5)
    w.add_circle(b, 0.3, category=(4, 0), mask=(0xFFFFFFFD, 0xFFFFFFFF))
    b = w.add_body(scenes.DYNAMIC, 30.0, 3.0)
    w.add_circle(b, 0.3, mask=(0xFFFFFFFE, 0xFFFFFFFF))
    # rounded polygons
    for i in range(4):
        b = w.add_body(scenes.DYNAMIC, 3.0 + 0.9 * i, 6.0 + rng.uniform(0.0, 1.0), angle=rng.uniform(-1.0, 1.0))
        pts = [(-0.3, -0.2), (0.35, -0.25), (0.1, 0.3)] if i % 2 == 0 else [(-0.3, -0.3), (0.3, -0.3), (0.3, 0.3),
                                                                            (-0.3, 0.3), (-0.4, 0.0)]
        w.add_polygon(b, pts, radius=0.04 + 0.03 * i)
    # revolute joints: motor wheel, spring pendulum, motor within limits, collideConnected
    anchor = w.add_body(scenes.STATIC, -2.0, 8.0)
    wheel = w.add_body(scenes.DYNAMIC, -2.0, 8.0)
    w.add_box(wheel, 1.0, 0.1)
    w.add_revolute(anchor, wheel, (0.0, 0.0), (0.0, 0.0), enable_motor=True, motor_speed=3.0, max_motor_torque=20.0)
    bob = w.add_body(scenes.DYNAMIC, 1.5, 8.0)
    w.add_circle(bob, 0.3)
    w.add_revolute(anchor, bob, (3.5, 0.0), (0.0, 0.0), enable_spring=True, hertz=2.0, damping=0.3,
                   enable_limit=True, lower=-0.8, upper=0.8)
    arm = w.add_body(scenes.DYNAMIC, -5.0, 8.0)
    w.add_box(arm, 1.0, 0.15)
    w.add_revolute(anchor, arm, (-2.0, 0.0), (1.0, 0.0), enable_limit=True, lower=-0.4, upper=0.6,
                   enable_motor=True, motor_speed=-2.0, max_motor_torque=5.0)
    a = w.add_body(scenes.DYNAMIC, 10.0, 6.0)
    w.add_box(a, 0.6, 0.3)
    c = w.add_body(scenes.DYNAMIC, 10.8, 6.0)
    w.add_box(c, 0.6, 0.3)
    w.add_revolute(a, c, (0.5, 0.0), (-0.3, 0.0), collide_connected=True)
    return w


def asleep_world():
    """Bodies created asleep (b2CreateBody with isAwake false, S/body.c:176-324): each
    gets a sleeping solver set and island of its own, keeps its place in the air and
    loses its initial velocity; a ball dropped on one wakes its set (b2LinkContact),
    another falls past them; a body created "asleep" that may not sleep, and a static
    body created asleep, are created awake."""
    w = scenes.World(scenes.world_header(sleep=True, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 30.0, 1.0)
    asleep = scenes.BF_ENABLE_SLEEP
    for i in range(3):
        b = w.add_body(scenes.DYNAMIC, -6.0 + 3.0 * i, 4.0, vx=2.0 * i, w=0.5, flags=asleep)
        w.add_box(b, 0.5, 0.5)
    b = w.add_body(scenes.DYNAMIC, -3.0, 9.0)
    w.add_circle(b, 0.4)
    b = w.add_body(scenes.DYNAMIC, 8.0, 6.0, vx=-1.0, flags=0)  # may not sleep: created awake
    w.add_capsule(b, (-0.4, 0.0), (0.4, 0.0), 0.2)
    s = w.add_body(scenes.STATIC, 12.0, 1.0, flags=asleep)
    w.add_box(s, 1.0, 1.0)
    b = w.add_body(scenes.DYNAMIC, 12.2, 4.0)
    w.add_box(b, 0.4, 0.4)
    return w


OP_DT = np.dtype([("step", "<i4"), ("op", "<i4"), ("target", "<i4"), ("flag", "<i4"), ("f", "<f4", (6,))])
# op codes of the harness's "ops" mode (oracle/ref_harness.c apply_op) and of
# tests/flat_util.py, which makes the same calls through the flat ABI
OPS = {"apply_force": 1, "apply_force_to_center": 2, "apply_torque": 3, "apply_linear_impulse": 4,
       "apply_linear_impulse_to_center": 5, "apply_angular_impulse": 6, "set_transform": 7,
       "set_linear_velocity": 8, "set_angular_velocity": 9, "set_mass_data": 10, "apply_mass_from_shapes": 11,
       "shape_set_density": 12, "shape_set_friction": 13, "shape_set_restitution": 14, "set_gravity_scale": 15,
       "set_linear_damping": 16, "set_angular_damping": 17, "set_fixed_rotation": 18, "set_bullet": 19,
       "set_awake": 20, "joint_wake_bodies": 23, "create_prismatic": 25, "prismatic_config": 26,
       "prismatic_target": 27, "create_weld": 28, "weld_config": 29, "create_wheel": 30, "wheel_config": 31,
       "create_distance": 32, "distance_config": 33, "distance_length": 34, "revolute_config": 35,
       "revolute_target": 36, "create_motor": 37, "motor_config": 38, "create_mouse": 39, "mouse_config": 40,
       "create_filter": 41, "destroy_joint": 42, "collide_connected": 43}


def ops_array(rows):
    """(step, name, target, flag, floats) rows -> OP_DT records in step order."""
    a = np.zeros(len(rows), OP_DT)
    for i, (step, name, target, flag, f) in enumerate(rows):
        a[i]["step"], a[i]["op"], a[i]["target"], a[i]["flag"] = step, OPS[name], target, flag
        a[i]["f"][:len(f)] = f
    return a[np.argsort(a["step"], kind="stable")]


def body_ops_world():
    """A scene for the body, shape and joint calls of the flat ABI: boxes, circles and a
    capsule on a ground box, a static post, a body that may not rotate, a hanging
    revolute chain, and (sleep on) a box created asleep."""
    w = scenes.World(scenes.world_header(sleep=True, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)                      # 0
    w.add_box(g, 30.0, 1.0)                                       # shape 0
    post = w.add_body(scenes.STATIC, 10.0, 1.0)                   # 1
    w.add_box(post, 0.3, 1.0)                                     # 1
    for i in range(4):                                            # bodies 2..5, shapes 2..5
        b = w.add_body(scenes.DYNAMIC, -6.0 + 2.0 * i, 0.5 + 0.01 * i)
        w.add_box(b, 0.5, 0.5, density=1.0 + 0.5 * i)
    b = w.add_body(scenes.DYNAMIC, 3.0, 2.0)                      # 6, shape 6, 7
    w.add_circle(b, 0.4)
    w.add_capsule(b, (0.0, 0.0), (0.8, 0.0), 0.2)
    b = w.add_body(scenes.DYNAMIC, 5.0, 0.5, flags=scenes.BF_FIXED_ROT | scenes.BF_AWAKE | scenes.BF_ENABLE_SLEEP)
    w.add_box(b, 0.5, 0.5)                                        # 7, shape 8
    b = w.add_body(scenes.DYNAMIC, -10.0, 3.0, flags=scenes.BF_ENABLE_SLEEP)  # 8: created asleep, shape 9
    w.add_box(b, 0.5, 0.5)
    anchor = w.add_body(scenes.STATIC, 15.0, 6.0)                 # 9, no shape
    prev, pa = anchor, (0.0, 0.0)
    for k in range(3):                                            # 10..12, shapes 10..12, joints 0..2
        b = w.add_body(scenes.DYNAMIC, 15.0, 6.0 - 0.2 - 0.4 * k)
        w.add_box(b, 0.05, 0.2)
        w.add_revolute(prev, b, pa, (0.0, 0.2), enable_limit=True, lower=-0.5, upper=0.5)
        prev, pa = b, (0.0, -0.2)
    return w


BODY_OPS = [
    (3, "apply_force", 2, 1, [50.0, 20.0, -6.2, 0.9]),
    (3, "apply_torque", 3, 1, [5.0]),
    (4, "apply_force_to_center", 4, 0, [-30.0, 0.0]),
    (5, "apply_linear_impulse", 5, 1, [2.0, 3.0, 0.2, 0.7]),
    (5, "apply_angular_impulse", 6, 1, [1.5]),
    (6, "apply_linear_impulse_to_center", 4, 1, [0.0, 1.0]),
    (8, "set_transform", 6, 0, [0.0, 4.0, 0.8660254, 0.5]),       # moves the proxy
    (8, "set_transform", 1, 0, [11.0, 1.0, 1.0, 0.0]),            # a static body's proxy moves
    (9, "set_transform", 3, 0, [-4.0, 0.52, 1.0, 0.0]),           # within its fat box: no move
    (10, "set_linear_velocity", 5, 0, [-1.0, 2.0]),
    (10, "set_angular_velocity", 2, 0, [3.0]),
    (10, "set_angular_velocity", 7, 0, [3.0]),                    # fixed rotation: ignored
    (12, "set_mass_data", 4, 0, [3.0, 0.7, 0.1, -0.05]),
    (14, "shape_set_density", 2, 1, [4.0]),
    (14, "shape_set_friction", 3, 0, [0.1]),
    (14, "shape_set_restitution", 4, 0, [0.5]),
    (15, "set_gravity_scale", 5, 0, [0.4]),
    (15, "set_linear_damping", 6, 0, [0.5]),
    (15, "set_angular_damping", 6, 0, [0.9]),
    (16, "set_fixed_rotation", 2, 1, []),
    (16, "set_bullet", 6, 1, []),
    (18, "apply_mass_from_shapes", 4, 0, []),
    (19, "apply_force", 8, 0, [0.0, 100.0, -10.0, 3.0]),           # asleep, no wake: nothing
    (20, "apply_linear_impulse_to_center", 8, 1, [30.0, 0.0]),     # wakes it
    (21, "apply_linear_impulse_to_center", 5, 0, [1000.0, 0.0]),   # beyond the speed limit
    (22, "set_fixed_rotation", 2, 0, []),
    (24, "set_awake", 7, 0, []),
    (26, "joint_wake_bodies", 1, 0, []),
    (28, "set_transform", 10, 0, [15.5, 5.8, 1.0, 0.0]),
]

def prismatic_world():
    """A scene for prismatic joints (S/prismatic_joint.c), created by API calls: boxes,
    a circle and a fixed-rotation box on a ground box, two boxes to be joined to each
    other, and loose boxes that fall on the jointed ones (sleep off)."""
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)                      # 0
    w.add_box(g, 30.0, 1.0)                                       # shape 0
    w.add_box(w.add_body(scenes.DYNAMIC, -6.0, 3.0), 0.5, 0.5)    # 1
    w.add_box(w.add_body(scenes.DYNAMIC, -2.0, 2.0, angle=0.3), 0.4, 0.6)  # 2
    w.add_box(w.add_body(scenes.DYNAMIC, 3.0, 1.0), 0.5, 0.3)     # 3
    w.add_box(w.add_body(scenes.DYNAMIC, 4.5, 1.0), 0.5, 0.3)     # 4
    w.add_circle(w.add_body(scenes.DYNAMIC, 8.0, 4.0), 0.5)       # 5
    fixed = scenes.BF_FIXED_ROT | scenes.BF_AWAKE | scenes.BF_ENABLE_SLEEP
    w.add_box(w.add_body(scenes.DYNAMIC, 11.0, 2.0, flags=fixed), 0.5, 0.5)  # 6
    for k in range(4):                                            # 7..10: fall onto the jointed bodies
        w.add_box(w.add_body(scenes.DYNAMIC, -6.2 + 4.7 * k, 6.0 + 0.5 * k, angle=0.2 * k), 0.3, 0.3)
    return w


PRISMATIC_OPS = [
    # joints 0..5 created before step 1 (flag: body B | collideConnected << 16)
    (1, "create_prismatic", 0, 1, [-6.0, 0.0, 0.0, -1.0, 0.0, 1.0]),             # 0: vertical slider
    (1, "prismatic_config", 0, 3, [-1.0, 2.0, 50.0, 1.0, 0.0, 0.0]),             # limit + motor
    (1, "create_prismatic", 0, 2, [-2.0, 1.0, 0.0, 0.0, 1.0, 1.0]),              # 1: axis (1, 1) normalized
    (1, "prismatic_config", 1, 4, [0.0, 0.0, 0.0, 0.0, 2.0, 0.5]),               # spring
    (1, "prismatic_target", 1, 0, [0.5]),
    (1, "create_prismatic", 3, 4, [0.5, 0.0, -0.5, 0.0, 1.0, 0.0]),              # 2: box to box
    (1, "prismatic_config", 2, 3, [0.5, 2.0, 20.0, -2.0, 0.0, 0.0]),             # limit + motor
    (1, "create_prismatic", 0, 5 | (1 << 16), [8.0, 1.0, 0.0, 0.0, 0.3, -1.0]),  # 3: collide connected
    (1, "prismatic_config", 3, 7, [-3.0, 1.5, 30.0, 0.5, 4.0, 0.7]),             # limit + motor + spring
    (1, "create_prismatic", 6, 0, [0.0, 0.0, 11.0, 1.0, 1.0, 0.0]),              # 4: fixed rotation on A
    (1, "prismatic_config", 4, 1, [-2.0, 2.0, 0.0, 0.0, 0.0, 0.0]),
    # later changes: target, motor reversal, limits moved, spring off, limit off
    (20, "prismatic_target", 1, 0, [-0.5]),
    (25, "prismatic_config", 0, 3, [-1.0, 2.0, 50.0, -1.5, 0.0, 0.0]),
    (30, "prismatic_config", 2, 3, [0.2, 1.0, 20.0, -2.0, 0.0, 0.0]),
    (35, "prismatic_config", 3, 3, [-3.0, 1.5, 30.0, 0.5, 4.0, 0.7]),
    (40, "prismatic_config", 4, 0, [-2.0, 2.0, 0.0, 0.0, 0.0, 0.0]),
]

def weld_wheel_world():
    """A scene for weld and wheel joints (S/weld_joint.c, S/wheel_joint.c), created by API
    calls: a box to weld to the ground, two boxes to weld to each other (one soft
    weld), a chassis box with two wheels, a box and a circle on a slope of boxes, and
    loose boxes that fall on them (sleep off)."""
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)                      # 0
    w.add_box(g, 30.0, 1.0)                                       # shape 0
    w.add_box(w.add_body(scenes.DYNAMIC, -8.0, 2.0), 0.5, 0.5)    # 1: welded to the ground
    w.add_box(w.add_body(scenes.DYNAMIC, -4.0, 3.0), 0.6, 0.2)    # 2
    w.add_box(w.add_body(scenes.DYNAMIC, -2.6, 3.2, angle=0.4), 0.6, 0.2)  # 3: soft weld to 2
    w.add_box(w.add_body(scenes.DYNAMIC, 4.0, 1.5), 1.5, 0.3)     # 4: chassis
    w.add_circle(w.add_body(scenes.DYNAMIC, 3.0, 0.9), 0.4)       # 5: rear wheel
    w.add_circle(w.add_body(scenes.DYNAMIC, 5.0, 0.9), 0.4)       # 6: front wheel
    w.add_box(w.add_body(scenes.DYNAMIC, 10.0, 1.0, angle=0.7), 0.4, 0.4)  # 7: welded with a reference angle
    for k in range(3):                                            # 8..10: fall onto the jointed bodies
        w.add_box(w.add_body(scenes.DYNAMIC, -7.5 + 6.0 * k, 6.0 + 0.5 * k, angle=0.3 * k), 0.3, 0.3)
    return w


WELD_WHEEL_OPS = [
    (1, "create_weld", 0, 1, [-8.0, 1.5, 0.0, -0.5, 0.0]),                   # 0: rigid weld to the ground
    (1, "create_weld", 2, 3, [0.7, 0.0, -0.7, 0.0, 0.4]),                    # 1: box to box
    (1, "weld_config", 1, 0, [5.0, 3.0, 0.7, 0.5]),                          #    soft in both senses
    (1, "create_wheel", 4, 5, [-1.0, -0.6, 0.0, 0.0, 0.0, 1.0]),             # 2: rear wheel, default spring
    (1, "wheel_config", 2, 7, [-0.02, 0.3, 30.0, -6.0, 4.0, 0.7]),           #    limit + motor + spring
    (1, "create_wheel", 4, 6 | (1 << 16), [1.0, -0.6, 0.0, 0.0, 0.3, 2.0]),  # 3: front wheel, axis normalized
    (1, "create_weld", 0, 7, [10.0, 2.0, 0.0, 0.0, 0.7]),                    # 4: reference angle
    (1, "weld_config", 4, 0, [0.0, 2.0, 0.0, 0.3]),                          #    angular spring only
    # later changes: motor reversed, spring off, a stiffer weld, the limits moved so
    # that the upper one acts, limit off
    (25, "wheel_config", 2, 7, [0.05, 0.3, 30.0, 4.0, 4.0, 0.7]),
    (35, "wheel_config", 3, 0, [0.0, 0.0, 0.0, 0.0, 1.0, 0.7]),
    (40, "weld_config", 1, 0, [20.0, 10.0, 1.0, 1.0]),
    (45, "wheel_config", 2, 7, [-0.3, -0.25, 30.0, 4.0, 4.0, 0.7]),
    (55, "wheel_config", 2, 6, [-0.3, -0.25, 30.0, 4.0, 4.0, 0.7]),
]

def distance_world():
    """A scene for distance joints (S/distance_joint.c), created by API calls: a pendulum
    box and a circle hanging from a static anchor, two boxes on the ground, a box
    hanging from a box, and loose boxes that fall on them (sleep off)."""
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)                      # 0
    w.add_box(g, 30.0, 1.0)                                       # shape 0
    anchor = w.add_body(scenes.STATIC, 0.0, 8.0)                  # 1, no shape
    w.add_box(w.add_body(scenes.DYNAMIC, 3.0, 7.0), 0.3, 0.3)     # 2: rigid pendulum
    w.add_circle(w.add_body(scenes.DYNAMIC, -2.0, 5.0), 0.4)      # 3: spring with a length range
    w.add_box(w.add_body(scenes.DYNAMIC, 6.0, 0.5), 0.5, 0.5)     # 4
    w.add_box(w.add_body(scenes.DYNAMIC, 8.0, 0.5), 0.5, 0.5)     # 5: motor along 4-5
    w.add_box(w.add_body(scenes.DYNAMIC, -8.0, 4.0), 0.6, 0.3)    # 6
    w.add_box(w.add_body(scenes.DYNAMIC, -8.5, 2.0, angle=0.5), 0.3, 0.3)  # 7: hangs from 6
    for k in range(3):                                            # 8..10
        w.add_box(w.add_body(scenes.DYNAMIC, -7.8 + 7.0 * k, 9.0 + 0.4 * k, angle=0.2 * k), 0.3, 0.3)
    return w


DISTANCE_OPS = [
    (1, "create_distance", 1, 2, [0.0, 0.0, 0.0, 0.2, 3.0]),                 # 0: rigid rope
    (1, "create_distance", 1, 3, [0.0, -0.5, 0.0, 0.0, 2.5]),                # 1
    (1, "distance_config", 1, 5, [1.5, 3.0, 0.0, 0.0, 3.0, 0.3]),            #    spring + range
    (1, "create_distance", 4, 5, [0.5, 0.0, -0.5, 0.0, 1.0]),                # 2
    (1, "distance_config", 2, 7, [0.8, 2.0, 40.0, 3.0, 0.0, 0.0]),           #    range + motor, spring without force
    (1, "create_distance", 6, 7 | (1 << 16), [-0.5, 0.0, 0.0, 0.3, 1.2]),    # 3: collide connected
    (1, "distance_config", 3, 5, [0.8, 1.6, 0.0, 0.0, 5.0, 0.9]),
    # later changes: new length, motor reversed, spring off (rigid), range tightened
    (20, "distance_length", 0, 0, [2.0]),
    (25, "distance_config", 2, 7, [0.8, 2.0, 40.0, -4.0, 0.0, 0.0]),
    (40, "distance_config", 1, 1, [1.5, 3.0, 0.0, 0.0, 3.0, 0.3]),
    (45, "distance_config", 3, 5, [1.0, 1.1, 0.0, 0.0, 5.0, 0.9]),
]

def motor_mouse_world():
    """A scene for motor, mouse and filter joints (S/motor_joint.c, S/mouse_joint.c,
    S/joint.c:480-505), created by API calls: a box driven by a motor joint towards an
    offset from the ground, a box carried by a motor joint on another box, a box and a
    circle dragged by mouse joints, two overlapping boxes kept from colliding by a filter
    joint, and loose boxes that fall on them (sleep off)."""
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)                      # 0
    w.add_box(g, 30.0, 1.0)                                       # shape 0
    w.add_box(w.add_body(scenes.DYNAMIC, -6.0, 0.5), 0.5, 0.5)    # 1: motor from the ground
    w.add_box(w.add_body(scenes.DYNAMIC, -2.0, 0.6), 1.0, 0.4)    # 2: carrier
    w.add_box(w.add_body(scenes.DYNAMIC, -2.0, 1.6, angle=0.3), 0.3, 0.3)  # 3: motor on 2
    w.add_box(w.add_body(scenes.DYNAMIC, 3.0, 0.5), 0.5, 0.5)     # 4: mouse
    w.add_circle(w.add_body(scenes.DYNAMIC, 6.0, 3.0), 0.4)       # 5: mouse, weak force
    w.add_box(w.add_body(scenes.DYNAMIC, 9.0, 0.5), 0.5, 0.5)     # 6
    w.add_box(w.add_body(scenes.DYNAMIC, 9.4, 1.2), 0.5, 0.5)     # 7: filtered against 6
    for k in range(3):                                            # 8..10
        w.add_box(w.add_body(scenes.DYNAMIC, -5.5 + 6.0 * k, 6.0 + 0.5 * k, angle=0.2 * k), 0.3, 0.3)
    return w


MOTOR_MOUSE_OPS = [
    (1, "create_motor", 0, 1, [0.5, 2.0, 0.8, 30.0, 5.0, 0.3]),               # 0
    (1, "create_motor", 2, 3, [0.0, 1.0, 0.0, 8.0, 1.0, 0.5]),                # 1
    (1, "create_mouse", 0, 4, [3.0, 0.5, 5.0, 0.7, 500.0]),                   # 2
    (1, "create_mouse", 0, 5 | (1 << 16), [6.0, 3.0, 2.0, 0.2, 3.0]),         # 3: weak force, clamped
    (1, "create_filter", 6, 7, []),                                           # 4
    # later: targets dragged, offsets and limits changed (negative force clamped to 0,
    # correction factor clamped to 1)
    (15, "mouse_config", 2, 0, [5.0, 3.0, 5.0, 0.7, 500.0]),
    (20, "motor_config", 0, 0, [-1.0, 3.0, -0.5, 30.0, 5.0, 0.3]),
    (25, "mouse_config", 3, 0, [2.0, 6.0, 8.0, 1.0, 40.0]),
    (30, "motor_config", 1, 0, [0.5, 1.0, 0.4, -2.0, 1.0, 1.5]),
    (40, "mouse_config", 2, 0, [1.0, 4.0, 1.0, 0.0, 50.0]),
]

# motor_mouse_world with joints destroyed: a mouse joint released and another one
# made (it takes the freed id), a motor and the filter joint destroyed (the filtered
# boxes then collide), a joint made after them
DESTROY_OPS = [
    (1, "create_motor", 0, 1, [0.5, 2.0, 0.8, 30.0, 5.0, 0.3]),               # 0
    (1, "create_motor", 2, 3, [0.0, 1.0, 0.0, 8.0, 1.0, 0.5]),                # 1
    (1, "create_mouse", 0, 4, [3.0, 0.5, 5.0, 0.7, 500.0]),                   # 2
    (1, "create_filter", 6, 7, []),                                           # 3
    (1, "create_mouse", 0, 5, [6.0, 3.0, 2.0, 0.2, 30.0]),                    # 4
    (10, "mouse_config", 2, 0, [5.0, 3.0, 5.0, 0.7, 500.0]),
    (20, "destroy_joint", 2, 0, []),                                          # mouse released
    (25, "create_mouse", 0, 4, [2.0, 4.0, 5.0, 0.7, 500.0]),                  # 5: takes id 2
    (30, "destroy_joint", 1, 0, []),                                          # motor on the carrier
    (35, "destroy_joint", 3, 0, []),                                          # filter: 6 and 7 collide
    (40, "create_motor", 2, 3, [0.0, 1.2, 0.0, 8.0, 1.0, 0.5]),              # 6: takes id 3
    (45, "destroy_joint", 0, 0, []),
]

# motor_mouse_world: a box pulled into its carrier by a motor joint without collision,
# then collideConnected turned on (the box's shape enters the move buffer and the two
# collide); a joint between distant boxes turned off and on again
COLLIDE_CONNECTED_OPS = [
    (1, "create_motor", 2, 3, [0.0, 1.0, 0.0, 8.0, 1.0, 0.5]),                # 0
    (1, "create_motor", 1, 4 | (1 << 16), [9.0, 0.0, 0.0, 0.0, 0.0, 0.0]),   # 1: no force
    (10, "motor_config", 0, 0, [0.0, 0.55, 0.0, 8.0, 1.0, 0.5]),
    (20, "collide_connected", 0, 1, []),
    (30, "collide_connected", 1, 0, []),
    (40, "collide_connected", 1, 1, []),
]

# the revolute chain of body_ops_world (joints 0..2): motor, limits, spring and target
# changed through the revolute joint calls
REVOLUTE_OPS = [
    (2, "revolute_config", 0, 2, [-0.5, 0.5, 5.0, 2.0, 0.0, 0.0]),    # motor on, limit off
    (10, "revolute_config", 1, 5, [-0.2, 0.3, 0.0, 0.0, 2.0, 0.5]),   # limit + spring
    (10, "revolute_target", 1, 0, [0.25]),
    (20, "revolute_config", 0, 3, [-0.3, 0.3, 5.0, -3.0, 0.0, 0.0]),  # motor reversed, limits moved
    (30, "revolute_config", 2, 4, [-0.5, 0.5, 0.0, 0.0, 4.0, 0.1]),   # limit off, spring on
    (35, "revolute_config", 1, 1, [-0.2, 0.3, 0.0, 0.0, 2.0, 0.5]),   # spring off
]

OPS_FIXTURES = {
    # name: (world factory, ops, dumped steps)
    "ops_bodies": (body_ops_world, BODY_OPS, list(range(0, 41))),
    "ops_prismatic": (prismatic_world, PRISMATIC_OPS, list(range(0, 61))),
    "ops_weld_wheel": (weld_wheel_world, WELD_WHEEL_OPS, list(range(0, 61))),
    "ops_distance": (distance_world, DISTANCE_OPS, list(range(0, 61))),
    "ops_revolute": (body_ops_world, REVOLUTE_OPS, list(range(0, 41))),
    "ops_motor_mouse": (motor_mouse_world, MOTOR_MOUSE_OPS, list(range(0, 61))),
    "ops_destroy_joint": (motor_mouse_world, DESTROY_OPS, list(range(0, 61))),
    "ops_collide_connected": (motor_mouse_world, COLLIDE_CONNECTED_OPS, list(range(0, 61))),
}


def make_ops(name: str):
    factory, rows, steps = OPS_FIXTURES[name]
    spec = scenes.to_bytes([factory()])
    ops = ops_array(rows)
    with tempfile.TemporaryDirectory() as td:
        sp, op, dp = os.path.join(td, "scene.bin"), os.path.join(td, "ops.bin"), os.path.join(td, "snaps.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        with open(op, "wb") as f:
            f.write(ops.tobytes())
        subprocess.run([HARNESS, "ops", sp, op, dp, "0", str(max(steps)), ",".join(map(str, steps))], check=True)
        with open(dp, "rb") as f:
            snaps = f.read()
        gp = os.path.join(td, "geoms.bin")
        subprocess.run([HARNESS, "geom", sp, gp, "0"], check=True)
        with open(gp, "rb") as f:
            geoms = f.read()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, spec=np.frombuffer(spec, np.uint8), snaps=np.frombuffer(snaps, np.uint8),
                        steps=np.array(steps, np.int32), geoms=np.frombuffer(geoms, np.uint8),
                        ops=np.frombuffer(ops.tobytes(), np.uint8))
    print(f"{name}: {len(ops)} calls, {len(snaps)} B snapshots -> {os.path.getsize(path)} B")


def query_world():
    """World queries (S/world.c:2050-2308, 2640-2747): every shape type on static,
    kinematic and dynamic bodies, rounded boxes (ray cast by shape cast), a body with
    two shapes, and shape categories 1, 2, 4 and one in the high word (some masks
    leave category 2 out), sleep off so that explosions push awake bodies only."""
    rng = scenes.SplitMix64(4242)
    w = scenes.World(scenes.world_header(sleep=False, continuous=True))
    g = w.add_body(scenes.STATIC, 0.0, -1.0)
    w.add_box(g, 30.0, 1.0)
    walls = w.add_body(scenes.STATIC, 0.0, 0.0)
    w.add_segment(walls, (-30.0, 0.0), (-30.0, 20.0))
    w.add_segment(walls, (30.0, 0.0), (30.0, 20.0), category=(2, 0))
    rb = w.add_body(scenes.STATIC, 12.0, 4.0, angle=0.3)
    w.add_box(rb, 1.0, 0.5, radius=0.2, category=(4, 0))
    sc = w.add_body(scenes.STATIC, -12.0, 6.0)
    w.add_circle(sc, 1.0, category=(2, 0))
    k1 = w.add_body(scenes.KINEMATIC, 0.0, 12.0, vx=1.0, w=0.5)
    w.add_capsule(k1, (-1.0, 0.0), (1.0, 0.0), 0.3, category=(2, 0))
    k2 = w.add_body(scenes.KINEMATIC, -5.0, 14.0, w=1.0)
    w.add_box(k2, 0.8, 0.4)
    cats = [(1, 0), (2, 0), (4, 0), (0, 1)]
    for i in range(36):
        b = w.add_body(scenes.DYNAMIC, rng.uniform(-9.0, 9.0), rng.uniform(1.0, 10.0),
                       angle=rng.uniform(-3.0, 3.0))
        _ = scenes._random_shape(w, rng, b)
        shape = list(w.shapes[-1])
        shape[7], shape[8] = cats[i % 4]
        if i % 5 == 0:
            shape[9] = 0xFFFFFFFD  # mask without category 2
        w.shapes[-1] = tuple(shape)
    for i in range(3):
        b = w.add_body(scenes.DYNAMIC, -6.0 + 6.0 * i, 16.0, angle=0.2 * i)
        w.add_box(b, 0.5, 0.3, radius=0.1)
    two = w.add_body(scenes.DYNAMIC, 24.0, 3.0)
    w.add_box(two, 0.5, 0.5)
    w.add_circle(two, 0.4, cx=0.8, cy=0.0)
    for r in range(4):
        n = 4 - r
        for i in range(n):
            b = w.add_body(scenes.DYNAMIC, -20.0 + (i - (n - 1) / 2.0), 0.5 + r)
            w.add_box(b, 0.5, 0.5)
    return w


def random_queries(snap, rng, explosions, n_rays=160, n_boxes=40):
    """Seeded rays (some from inside bodies, some of zero length), AABB boxes and the
    given explosions, with query filters over categories 1, 2, 4 and the high word."""
    from forge2d_amd import make_queries, Q_RAY_CLOSEST, Q_RAY_ALL, Q_OVERLAP_AABB, Q_EXPLODE

    filters = [(1, 0, 0xFFFFFFFF, 0xFFFFFFFF)] * 7 + [(1, 0, 2, 0), (1, 0, 4, 1), (2, 0, 0xFFFFFFFF, 0xFFFFFFFF)]
    pos = np.stack([snap.bodies["px"], snap.bodies["py"]], 1)
    rows = []
    for i in range(n_rays):
        if i % 8 == 0:
            o = pos[rng.randint(0, len(pos) - 1)]
            ox, oy = float(o[0]), float(o[1])
        else:
            ox, oy = rng.uniform(-32.0, 32.0), rng.uniform(-3.0, 22.0)
        a = rng.uniform(-math.pi, math.pi)
        ln = 0.0 if i % 23 == 0 else rng.uniform(0.5, 40.0)
        tx, ty = ln * math.cos(a), ln * math.sin(a)
        kind = Q_RAY_CLOSEST if i % 2 == 0 else Q_RAY_ALL
        rows.append((0, kind, [ox, oy, tx, ty], filters[rng.randint(0, len(filters) - 1)]))
    for i in range(n_boxes):
        x, y = rng.uniform(-32.0, 32.0), rng.uniform(-3.0, 22.0)
        rows.append((0, Q_OVERLAP_AABB, [x, y, x + rng.uniform(0.1, 12.0), y + rng.uniform(0.1, 8.0)],
                     filters[rng.randint(0, len(filters) - 1)]))
    for e in explosions:
        rows.append((0, Q_EXPLODE, e[:5], (0, 0, e[5], e[6])))
    return make_queries(rows)


QUERY_FIXTURES = {
    # name: (world factory, steps before the queries, query seed, explosions
    #        (x, y, radius, falloff, impulse per length, mask lo, mask hi))
    "queries": (query_world, 90, 7,
                [(0.0, 1.0, 3.0, 0.0, 2.0, 0xFFFFFFFF, 0xFFFFFFFF), (-4.0, 0.5, 2.0, 2.0, 1.5, 0xFFFFFFFF, 0xFFFFFFFF),
                 (5.0, 1.0, 0.0, 3.0, 1.0, 0xFFFFFFFF, 0xFFFFFFFF), (0.0, 2.0, 6.0, 1.0, 0.7, 0xFFFFFFFD, 0xFFFFFFFF),
                 (-20.0, 2.0, 1.5, 1.0, 3.0, 0xFFFFFFFF, 0xFFFFFFFF)]),
    # after 200 steps the right pyramid sleeps as one island: the explosion wakes it
    "queries_sleep": (sleep_world, 200, 8, [(10.0, 1.0, 1.5, 1.0, 2.0, 0xFFFFFFFF, 0xFFFFFFFF)]),
    # two boxes sleep apart (two islands): one explosion wakes both, in the reference's
    # tree order (the restatements wake them in shape-id order)
    "queries_wake2": (sleep_world, 200, 9, [(-0.3, 0.5, 0.6, 0.5, 1.0, 0xFFFFFFFF, 0xFFFFFFFF)]),
}


def make_queries_fixture(name: str):
    """Query fixture: the reference's results and its snapshots before and after."""
    from forge2d_amd import snapshot as S

    factory, steps, seed, explosions = QUERY_FIXTURES[name]
    spec = scenes.to_bytes([factory()])
    with tempfile.TemporaryDirectory() as td:
        sp, dp = os.path.join(td, "scene.bin"), os.path.join(td, "d.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        subprocess.run([HARNESS, "dump", sp, dp, "0", str(steps), str(steps)], check=True)
        with open(dp, "rb") as f:
            snap = S.parse_one(f.read())
        queries = random_queries(snap, scenes.SplitMix64(seed), explosions)
        qp, rp = os.path.join(td, "q.bin"), os.path.join(td, "r.bin")
        with open(qp, "wb") as f:
            f.write(queries.tobytes())
        subprocess.run([HARNESS, "query", sp, qp, rp, "0", str(steps)], check=True)
        with open(rp, "rb") as f:
            results = f.read()
        with open(rp + ".snap", "rb") as f:
            snaps = f.read()
        gp = os.path.join(td, "geoms.bin")
        subprocess.run([HARNESS, "geom", sp, gp, "0"], check=True)
        with open(gp, "rb") as f:
            geoms = f.read()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, spec=np.frombuffer(spec, np.uint8), snaps=np.frombuffer(snaps, np.uint8),
                        steps=np.array([steps], np.int32), geoms=np.frombuffer(geoms, np.uint8),
                        queries=np.frombuffer(queries.tobytes(), np.uint8), results=np.frombuffer(results, np.uint8))
    print(f"{name}: {len(queries)} queries, {len(results)} B results -> {os.path.getsize(path)} B")


def parse_results(buf: bytes, n: int):
    """The harness's query results: per query an int32 count and that many 24-byte hits."""
    from forge2d_amd import HIT_DTYPE

    out, off = [], 0
    for _ in range(n):
        c = int(np.frombuffer(buf, np.int32, 1, off)[0])
        off += 4
        out.append(np.frombuffer(buf, HIT_DTYPE, c, off).copy())
        off += 24 * c
    assert off == len(buf)
    return out


EVENT_STEPS = list(range(1, 121))
# fixtures that also record the reference's contact and body events per step
EVENT_FIXTURES = {"events": EVENT_STEPS, "sleep": list(range(1, 261))}


FIXTURES = {
    # name: (world factory, dumped steps)
    "pyramid12": (lambda: scenes.pyramid_world(12), [0, 1, 2, 3, 4, 5, 60, 61, 62]),
    "pyramid40": (lambda: scenes.pyramid_world(40), [0, 1, 2, 3, 254, 255, 256]),
    "mixed": (lambda: scenes.container_world(150, 12.0, 24.0, 20251202, continuous=False, chains=2,
                                             drop_band=(0.1, 0.6)),
              [0, 1, 2, 3, 4, 5, 6, 30, 31, 32, 60, 61, 62, 120, 121, 122]),
    "overflow": (overflow_world, [0, 1, 2, 3, 4, 20, 21, 22]),
    "bounce": (bounce_world, [0, 40, 41, 42, 45, 46, 47, 50, 51, 52, 60, 61, 62, 80, 81, 82]),
    "ccd": (ccd_world, list(range(0, 13)) + [30, 31, 32]),
    "events": (events_world, list(range(0, 121))),
    "sleep": (sleep_world, list(range(0, 261))),
    "features": (features_world, list(range(0, 7)) + [30, 31, 32, 60, 61, 62, 120, 121, 122]),
    "asleep": (asleep_world, list(range(0, 61)) + [120, 121, 122]),
    "mixed_ccd": (lambda: scenes.container_world(200, 15.0, 40.0, 20261202, continuous=True, chains=3,
                                                 drop_band=(0.2, 1.0)),
                  [0, 1, 2, 3, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 60, 61, 62]),
}


# large worlds whose snapshots are too big to keep: the reference's state digests
# (tests.parity_util.state_digest) after the listed steps, from the initial state
DIGEST_FIXTURES = {
    # config 2: one 20,000-body world; its first steps and a window after 300 steps
    "config2": (lambda: scenes.config2()[0], [1, 2, 3, 300, 301, 600]),
}


def config2_outside(px, py) -> list:
    """Dynamic bodies whose centre left config 2's container [-100, 100] x [0, 200]."""
    px, py = np.asarray(px)[1:], np.asarray(py)[1:]
    return (np.nonzero((np.abs(px) >= 100.0) | (py <= 0.0) | (py >= 200.0))[0] + 1).tolist()


def make_digest(name: str):
    import hashlib
    import json

    from forge2d_amd import snapshot as S
    from tests import parity_util as P

    factory, steps = DIGEST_FIXTURES[name]
    spec = scenes.to_bytes([factory()])
    out = {"spec_sha256": hashlib.sha256(spec).hexdigest(), "steps": steps, "digests": []}
    with tempfile.TemporaryDirectory() as td:
        sp, dp = os.path.join(td, "scene.bin"), os.path.join(td, "snaps.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        subprocess.run([HARNESS, "dump", sp, dp, "0", str(max(steps)), ",".join(map(str, steps))], check=True)
        with open(dp, "rb") as f:
            snaps = S.parse_all(f.read())
    for sn in snaps:
        d = P.state_digest(sn)
        if name == "config2":
            d["outside_container"] = config2_outside(sn.bodies["px"], sn.bodies["py"])
        out["digests"].append(d)
    path = os.path.join(HERE, name + "_digest.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(f"{name}: digests of steps {steps} -> {os.path.getsize(path)} B "
          f"(live contacts {[d['live_contacts'] for d in out['digests']]}, "
          f"fast bodies {[d['fast_bodies'] for d in out['digests']]})")


def make(name: str):
    factory, steps = FIXTURES[name]
    world = factory()
    spec = scenes.to_bytes([world])
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        out = os.path.join(td, "snaps.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        subprocess.run([HARNESS, "dump", sp, out, "0", str(max(steps)), ",".join(map(str, steps))], check=True)
        with open(out, "rb") as f:
            snaps = f.read()
        gp = os.path.join(td, "geoms.bin")
        subprocess.run([HARNESS, "geom", sp, gp, "0"], check=True)
        with open(gp, "rb") as f:
            geoms = f.read()
        extra = {}
        if name in EVENT_FIXTURES:
            ev_steps = EVENT_FIXTURES[name]
            ep = os.path.join(td, "events.bin")
            subprocess.run([HARNESS, "events", sp, ep, "0", str(max(ev_steps)), ",".join(map(str, ev_steps))],
                           check=True)
            with open(ep, "rb") as f:
                extra["events"] = np.frombuffer(f.read(), np.uint8)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, spec=np.frombuffer(spec, np.uint8), snaps=np.frombuffer(snaps, np.uint8),
                        steps=np.array(steps, np.int32), geoms=np.frombuffer(geoms, np.uint8), **extra)
    print(f"{name}: {len(spec)} B spec, {len(snaps)} B snapshots -> {os.path.getsize(path)} B")


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    for n in sys.argv[1:] or list(FIXTURES) + list(QUERY_FIXTURES) + list(DIGEST_FIXTURES) + list(OPS_FIXTURES):
        if n in DIGEST_FIXTURES:
            make_digest(n)
        elif n in OPS_FIXTURES:
            make_ops(n)
        elif n in QUERY_FIXTURES:
            make_queries_fixture(n)
        else:
            make(n)
