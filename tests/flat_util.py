"""Scene specs through the flat per-world ABI (test infrastructure): a world built
with f2d_create_world / f2d_create_body / f2d_create_*_shape /
f2d_create_revolute_joint call by call, as a forge2d binding would build it, and the
API calls of an ops fixture (tests/golden/make_golden.py OPS, prismatic joints
included) replayed through the f2d_* functions."""
from __future__ import annotations

import ctypes

import numpy as np

import forge2d_amd as F
from forge2d_amd import scenes
from forge2d_amd import snapshot as S


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


class FlatWorld:
    """A world of one scene spec built through the flat ABI; ids as (index1, wg)."""

    def __init__(self, spec: bytes):
        L = F.lib()
        (h, bodies, shapes, joints), = scenes.parse(spec)
        flags = int(h["flags"])
        self.w = L.f2d_create_world(float(h["gx"]), float(h["gy"]), float(h["restitution_threshold"]),
                                    float(h["hit_event_threshold"]), float(h["contact_hertz"]),
                                    float(h["contact_damping_ratio"]), float(h["max_contact_push_speed"]),
                                    float(h["maximum_linear_speed"]), int(flags & scenes.WF_SLEEP != 0),
                                    int(flags & scenes.WF_CONTINUOUS != 0))
        out = (ctypes.c_int32 * 2)()
        self.bodies, self.shapes, self.joints = [], [], []
        for b in bodies:
            bf = int(b["flags"])
            L.f2d_create_body(self.w, int(b["type"]), float(b["px"]), float(b["py"]), float(b["qc"]), float(b["qs"]),
                              float(b["vx"]), float(b["vy"]), float(b["w"]), float(b["lin_damp"]),
                              float(b["ang_damp"]), float(b["gravity_scale"]), float(b["sleep_threshold"]), None,
                              int(bf & scenes.BF_ENABLE_SLEEP != 0), int(bf & scenes.BF_AWAKE != 0),
                              int(bf & scenes.BF_FIXED_ROT != 0), int(bf & scenes.BF_BULLET != 0), 1,
                              int(bf & scenes.BF_FAST_ROT != 0), out)
            assert out[0] > 0
            self.bodies.append((out[0], ctypes.c_uint32(out[1]).value))
        for s in shapes:
            d = ShapeDef()
            d.friction, d.restitution = float(s["friction"]), float(s["restitution"])
            d.rolling_resistance, d.tangent_speed = float(s["rolling"]), float(s["tangent_speed"])
            d.density = float(s["density"])
            d.category_lo, d.category_hi = int(s["cat_lo"]), int(s["cat_hi"])
            d.mask_lo, d.mask_hi, d.group_index = int(s["mask_lo"]), int(s["mask_hi"]), int(s["group"])
            sf = int(s["flags"])
            d.enable_contact_events = int(sf & scenes.SF_CONTACT_EVENTS != 0)
            d.enable_hit_events = int(sf & scenes.SF_HIT_EVENTS != 0)
            d.invoke_contact_creation = int(sf & scenes.SF_INVOKE_CONTACT != 0)
            d.enable_pre_solve_events = int(sf & scenes.SF_PRESOLVE != 0)
            d.update_body_mass = int(sf & scenes.SF_UPDATE_MASS != 0)
            d.is_sensor = int(sf & scenes.SF_SENSOR != 0)
            bi, bwg = self.bodies[int(s["body"])]
            k, v, r = int(s["kind"]), [float(x) for x in s["data"]], float(s["radius"])
            ref = ctypes.byref(d)
            if k == scenes.K_CIRCLE:
                L.f2d_create_circle_shape(bi, bwg, ref, v[0], v[1], r, out)
            elif k == scenes.K_CAPSULE:
                L.f2d_create_capsule_shape(bi, bwg, ref, v[0], v[1], v[2], v[3], r, out)
            elif k == scenes.K_SEGMENT:
                L.f2d_create_segment_shape(bi, bwg, ref, v[0], v[1], v[2], v[3], out)
            elif k == scenes.K_POLY:
                n = int(s["npts"])
                pts = (ctypes.c_float * (2 * n))(*v[:2 * n])
                assert L.f2d_create_polygon_shape(bi, bwg, ref, pts, n, r, out) == 1
            else:
                L.f2d_create_box_shape(bi, bwg, ref, v[0], v[1], v[2], v[3], v[4], v[5], r, out)
            assert out[0] > 0
            self.shapes.append((out[0], ctypes.c_uint32(out[1]).value))
        for j in joints:
            a, b = self.bodies[int(j["body_a"])], self.bodies[int(j["body_b"])]
            L.f2d_create_revolute_joint(self.w, a[0], a[1], b[0], b[1], float(j["lax"]), float(j["lay"]),
                                        float(j["lbx"]), float(j["lby"]), float(j["ref_angle"]), 0.0,
                                        int(j["enable_spring"]), float(j["hertz"]), float(j["damping"]),
                                        int(j["enable_limit"]), float(j["lower"]), float(j["upper"]),
                                        int(j["enable_motor"]), float(j["max_motor_torque"]),
                                        float(j["motor_speed"]), float(j["draw_size"]), int(j["collide_connected"]),
                                        out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))

    def step(self, dt=1.0 / 60.0, substeps=4):
        F.lib().f2d_world_step(self.w, dt, substeps)

    def snapshot_bytes(self) -> bytes:
        L = F.lib()
        b = L.f2d_world_batch(self.w)
        n = L.f2dg_export_world(b, 0, None, 0)
        buf = ctypes.create_string_buffer(n)
        assert L.f2dg_export_world(b, 0, buf, n) == n
        return buf.raw

    def snapshot(self) -> S.Snapshot:
        return S.parse_one(self.snapshot_bytes())

    def close(self):
        F.lib().f2d_destroy_world(self.w)

    def apply(self, op):
        """One ops-fixture record (OP_DT) through the flat ABI."""
        L = F.lib()
        code, t, flag = int(op["op"]), int(op["target"]), int(op["flag"])
        f = [float(x) for x in op["f"]]
        bi, bwg = self.bodies[t] if code not in (12, 13, 14, 23, 26, 27, 29, 31, 33, 34, 35, 36, 38, 40, 42, 43) \
            else (0, 0)
        if code == 1:
            L.f2d_body_apply_force(bi, bwg, f[0], f[1], f[2], f[3], flag)
        elif code == 2:
            L.f2d_body_apply_force_to_center(bi, bwg, f[0], f[1], flag)
        elif code == 3:
            L.f2d_body_apply_torque(bi, bwg, f[0], flag)
        elif code == 4:
            L.f2d_body_apply_linear_impulse(bi, bwg, f[0], f[1], f[2], f[3], flag)
        elif code == 5:
            L.f2d_body_apply_linear_impulse_to_center(bi, bwg, f[0], f[1], flag)
        elif code == 6:
            L.f2d_body_apply_angular_impulse(bi, bwg, f[0], flag)
        elif code == 7:
            L.f2d_body_set_transform(bi, bwg, f[0], f[1], f[2], f[3])
        elif code == 8:
            L.f2d_body_set_linear_velocity(bi, bwg, f[0], f[1])
        elif code == 9:
            L.f2d_body_set_angular_velocity(bi, bwg, f[0])
        elif code == 10:
            L.f2d_body_set_mass_data(bi, bwg, f[0], f[1], f[2], f[3])
        elif code == 11:
            L.f2d_body_apply_mass_from_shapes(bi, bwg)
        elif code == 12:
            L.f2d_shape_set_density(*self.shapes[t], f[0], flag)
        elif code == 13:
            L.f2d_shape_set_friction(*self.shapes[t], f[0])
        elif code == 14:
            L.f2d_shape_set_restitution(*self.shapes[t], f[0])
        elif code == 15:
            L.f2d_body_set_gravity_scale(bi, bwg, f[0])
        elif code == 16:
            L.f2d_body_set_linear_damping(bi, bwg, f[0])
        elif code == 17:
            L.f2d_body_set_angular_damping(bi, bwg, f[0])
        elif code == 18:
            L.f2d_body_set_fixed_rotation(bi, bwg, flag)
        elif code == 19:
            L.f2d_body_set_bullet(bi, bwg, flag)
        elif code == 20:
            L.f2d_body_set_awake(bi, bwg, flag)
        elif code == 23:
            L.f2d_joint_wake_bodies(*self.joints[t])
        elif code == 25:
            # a prismatic joint from the default def with anchors, axis and collideConnected
            bb = self.bodies[flag & 0xFFFF]
            out = (ctypes.c_int32 * 2)()
            L.f2d_create_prismatic_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4], f[5], 0.0,
                                         0.0, 0, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0, 0.0, int(flag >> 16 != 0), out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))
        elif code == 26:
            j = self.joints[t]
            L.f2d_prismatic_joint_set_limits(*j, f[0], f[1])
            L.f2d_prismatic_joint_enable_limit(*j, int(flag & 1 != 0))
            L.f2d_prismatic_joint_set_max_motor_force(*j, f[2])
            L.f2d_prismatic_joint_set_motor_speed(*j, f[3])
            L.f2d_prismatic_joint_enable_motor(*j, int(flag & 2 != 0))
            L.f2d_prismatic_joint_set_spring_hertz(*j, f[4])
            L.f2d_prismatic_joint_set_spring_damping_ratio(*j, f[5])
            L.f2d_prismatic_joint_enable_spring(*j, int(flag & 4 != 0))
        elif code == 27:
            L.f2d_prismatic_joint_set_target_translation(*self.joints[t], f[0])
        elif code == 28:
            # a weld joint from the default def with anchors, reference angle and collideConnected
            bb = self.bodies[flag & 0xFFFF]
            out = (ctypes.c_int32 * 2)()
            L.f2d_create_weld_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4], 0.0, 0.0, 0.0, 0.0,
                                    int(flag >> 16 != 0), out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))
        elif code == 29:
            j = self.joints[t]
            L.f2d_weld_joint_set_linear_hertz(*j, f[0])
            L.f2d_weld_joint_set_angular_hertz(*j, f[1])
            L.f2d_weld_joint_set_linear_damping_ratio(*j, f[2])
            L.f2d_weld_joint_set_angular_damping_ratio(*j, f[3])
        elif code == 30:
            # a wheel joint from the default def (spring on, 1 Hz, damping 0.7)
            bb = self.bodies[flag & 0xFFFF]
            out = (ctypes.c_int32 * 2)()
            L.f2d_create_wheel_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4], f[5], 1, 1.0, 0.7,
                                     0, 0.0, 0.0, 0, 0.0, 0.0, int(flag >> 16 != 0), out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))
        elif code == 31:
            j = self.joints[t]
            L.f2d_wheel_joint_set_limits(*j, f[0], f[1])
            L.f2d_wheel_joint_enable_limit(*j, int(flag & 1 != 0))
            L.f2d_wheel_joint_set_max_motor_torque(*j, f[2])
            L.f2d_wheel_joint_set_motor_speed(*j, f[3])
            L.f2d_wheel_joint_enable_motor(*j, int(flag & 2 != 0))
            L.f2d_wheel_joint_set_spring_hertz(*j, f[4])
            L.f2d_wheel_joint_set_spring_damping_ratio(*j, f[5])
            L.f2d_wheel_joint_enable_spring(*j, int(flag & 4 != 0))
        elif code == 32:
            # a distance joint from the default def (range [slop, huge], spring, limit and motor off)
            bb = self.bodies[flag & 0xFFFF]
            out = (ctypes.c_int32 * 2)()
            L.f2d_create_distance_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4], 0, 0.0, 0.0, 0,
                                        0.0, 100000.0, 0, 0.0, 0.0, int(flag >> 16 != 0), out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))
        elif code == 33:
            j = self.joints[t]
            L.f2d_distance_joint_set_length_range(*j, f[0], f[1])
            L.f2d_distance_joint_enable_limit(*j, int(flag & 1 != 0))
            L.f2d_distance_joint_set_max_motor_force(*j, f[2])
            L.f2d_distance_joint_set_motor_speed(*j, f[3])
            L.f2d_distance_joint_enable_motor(*j, int(flag & 2 != 0))
            L.f2d_distance_joint_set_spring_hertz(*j, f[4])
            L.f2d_distance_joint_set_spring_damping_ratio(*j, f[5])
            L.f2d_distance_joint_enable_spring(*j, int(flag & 4 != 0))
        elif code == 34:
            L.f2d_distance_joint_set_length(*self.joints[t], f[0])
        elif code == 35:
            j = self.joints[t]
            L.f2d_revolute_joint_set_limits(*j, f[0], f[1])
            L.f2d_revolute_joint_enable_limit(*j, int(flag & 1 != 0))
            L.f2d_revolute_joint_set_max_motor_torque(*j, f[2])
            L.f2d_revolute_joint_set_motor_speed(*j, f[3])
            L.f2d_revolute_joint_enable_motor(*j, int(flag & 2 != 0))
            L.f2d_revolute_joint_set_spring_hertz(*j, f[4])
            L.f2d_revolute_joint_set_spring_damping_ratio(*j, f[5])
            L.f2d_revolute_joint_enable_spring(*j, int(flag & 4 != 0))
        elif code == 36:
            L.f2d_revolute_joint_set_target_angle(*self.joints[t], f[0])
        elif code in (37, 39, 41):
            bb = self.bodies[flag & 0xFFFF]
            out = (ctypes.c_int32 * 2)()
            if code == 37:
                L.f2d_create_motor_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4], f[5],
                                         int(flag >> 16 != 0), out)
            elif code == 39:
                L.f2d_create_mouse_joint(self.w, bi, bwg, bb[0], bb[1], f[0], f[1], f[2], f[3], f[4],
                                         int(flag >> 16 != 0), out)
            else:
                L.f2d_create_filter_joint(self.w, bi, bwg, bb[0], bb[1], out)
            assert out[0] > 0
            self.joints.append((out[0], ctypes.c_uint32(out[1]).value))
        elif code == 42:
            L.f2d_destroy_joint(*self.joints[t])
        elif code == 43:
            L.f2d_joint_set_collide_connected(*self.joints[t], int(flag))
        elif code == 38:
            j = self.joints[t]
            L.f2d_motor_joint_set_linear_offset(*j, f[0], f[1])
            L.f2d_motor_joint_set_angular_offset(*j, f[2])
            L.f2d_motor_joint_set_max_force(*j, f[3])
            L.f2d_motor_joint_set_max_torque(*j, f[4])
            L.f2d_motor_joint_set_correction_factor(*j, f[5])
        elif code == 40:
            j = self.joints[t]
            L.f2d_mouse_joint_set_target(*j, f[0], f[1])
            L.f2d_mouse_joint_set_spring_hertz(*j, f[2])
            L.f2d_mouse_joint_set_spring_damping_ratio(*j, f[3])
            L.f2d_mouse_joint_set_max_force(*j, f[4])
        else:
            raise ValueError(f"op {code}")


def load_ops_fixture(name: str, golden: str):
    import os
    z = np.load(os.path.join(golden, name + ".npz"))
    from tests.golden.make_golden import OP_DT
    spec = z["spec"].tobytes()
    snaps_raw = z["snaps"].tobytes()
    snaps = S.parse_all(snaps_raw)
    ops = np.frombuffer(z["ops"].tobytes(), OP_DT)
    return spec, snaps, ops
