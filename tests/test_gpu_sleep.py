"""Sleep through the flat per-world ABI, replaying the reference's own sleep
assertions and the body API's wake/sleep semantics:

* F/test/ffi/smoke_test.dart:13-46 — a 1x1 box dropped from y=10 onto a 50x1
  ground box (sleep on, as b2DefaultWorldDef has it) rests at (0, 0.5) +- 0.01
  after 200 steps and is asleep;
* F/test/api/events_test.dart:134-148 — a box resting on the ground reports a body
  move event with fellAsleep within 300 steps;
* F/test/api/body_test.dart:50-66 — a static body is never in the awake set; the
  default sleep threshold is 0.05 and sleep is enabled;
* b2Body_SetAwake / b2WakeBody / b2Body_EnableSleep / b2World_EnableSleeping
  (S/body.c:332-344, 1483-1552; S/world.c:1710-1737): waking resets the sleep
  timer, an impulse with wake moves a sleeping body, a body with sleep disabled
  never sleeps, disabling sleep on the world wakes every sleeping body.
"""
import ctypes

import numpy as np
import pytest

import forge2d_amd as F
from tests.test_gpu_api import body, box, pos, vel

pytestmark = pytest.mark.gpu


def sleepy_world():
    L = F.lib()
    w = L.f2d_create_world(0.0, -10.0, 1.0, 1.0, 30.0, 10.0, 3.0, 400.0, 1, 1)
    assert L.f2d_world_is_sleeping_enabled(w) == 1
    ground = body(w, 0, 0.0, -1.0)
    box(ground, 50.0, 1.0)
    return w, ground


def move_events(w):
    L = F.lib()
    n = L.f2d_world_get_body_event_count(w)
    out = np.zeros(7 * max(n, 1), np.float64)
    if n:
        L.f2d_world_get_body_events(w, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out[:7 * n].reshape(n, 7)


def test_falling_box_rests_and_sleeps(gpu):
    L = F.lib()
    w, ground = sleepy_world()
    assert L.f2d_body_is_awake(ground[0], ground[1]) == 0  # static: never in the awake set
    b = body(w, 2, 0.0, 10.0)
    box(b, 0.5, 0.5)
    assert L.f2d_body_is_awake(b[0], b[1]) == 1
    assert L.f2d_body_is_sleep_enabled(b[0], b[1]) == 1
    assert abs(L.f2d_body_get_sleep_threshold(b[0], b[1]) - 0.05) < 1e-6
    for _ in range(200):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    x, y = pos(b)
    assert abs(x) < 0.01 and abs(y - 0.5) < 0.01, (x, y)
    assert L.f2d_body_is_awake(b[0], b[1]) == 0
    assert L.f2d_world_get_body_event_count(w) == 0  # nothing awake, nothing moved
    L.f2d_destroy_world(w)


def test_fell_asleep_event_then_wake_and_sleep_api(gpu):
    L = F.lib()
    w, _ = sleepy_world()
    b = body(w, 2, 0.0, 0.6)
    box(b, 0.5, 0.5)
    fell = -1
    for i in range(300):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
        ev = move_events(w)
        if len(ev) and ev[:, 6].any():
            fell = i
            assert ev[0, 0] == b[0]
            break
    assert fell >= 30, fell  # the sleep timer needs 0.5 s of rest
    assert L.f2d_body_is_awake(b[0], b[1]) == 0
    assert vel(b) == (0.0, 0.0)
    # waking resets the sleep timer: awake for at least another 0.5 s
    L.f2d_body_set_awake(b[0], b[1], 1)
    assert L.f2d_body_is_awake(b[0], b[1]) == 1
    for _ in range(25):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
        assert L.f2d_body_is_awake(b[0], b[1]) == 1
        assert len(move_events(w)) == 1 and not move_events(w)[0, 6]
    # put to sleep at once, then an impulse with wake moves it
    L.f2d_body_set_awake(b[0], b[1], 0)
    assert L.f2d_body_is_awake(b[0], b[1]) == 0
    L.f2d_body_apply_linear_impulse_to_center(b[0], b[1], 0.0, 0.0, 0)
    assert L.f2d_body_is_awake(b[0], b[1]) == 0  # no wake requested
    m = L.f2d_body_get_mass(b[0], b[1])
    L.f2d_body_apply_linear_impulse_to_center(b[0], b[1], 2.0 * m, 0.0, 1)
    assert L.f2d_body_is_awake(b[0], b[1]) == 1
    assert abs(vel(b)[0] - 2.0) < 1e-5
    # a body with sleep disabled stays awake
    L.f2d_body_enable_sleep(b[0], b[1], 0)
    assert L.f2d_body_is_sleep_enabled(b[0], b[1]) == 0
    for _ in range(120):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    assert L.f2d_body_is_awake(b[0], b[1]) == 1
    L.f2d_body_enable_sleep(b[0], b[1], 1)
    for _ in range(120):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
    assert L.f2d_body_is_awake(b[0], b[1]) == 0
    # disabling sleep on the world wakes every sleeping body
    L.f2d_world_enable_sleeping(w, 0)
    assert L.f2d_world_is_sleeping_enabled(w) == 0
    assert L.f2d_body_is_awake(b[0], b[1]) == 1
    L.f2d_destroy_world(w)
