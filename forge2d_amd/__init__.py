"""MI355X-native Box2D v3 world step (forge2d ``World.step``) — Python host binding.

The product is ``libforge2d_amd.so`` (HIP kernels for gfx950 behind a plain
C-ABI, ``include/forge2d_amd.h``). This module loads it with ctypes and offers
a thin batch wrapper used by the tests, the benchmark and the serving loop.
There is no CPU fallback: if the library is missing or cannot load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes, snapshot  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libforge2d_amd.so")

PH_PAIRS, PH_COLLIDE, PH_STATE, PH_SOLVE, PH_FINALIZE = 1, 2, 4, 8, 16
PH_ALL = 31
PH_CONTINUOUS, PH_NO_CONTINUOUS = 64, 128

_lib = None

# (name, restype, argtypes) of every exported symbol declared in include/forge2d_amd.h
_VP, _I, _L, _F, _U = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float, ctypes.c_uint32
_PI = ctypes.POINTER(ctypes.c_int32)
_PF = ctypes.POINTER(ctypes.c_float)
_PD = ctypes.POINTER(ctypes.c_double)
SIGNATURES = [
    ("f2dg_create_from_spec", _VP, [_VP, _L]),
    ("f2dg_destroy", None, [_VP]),
    ("f2dg_world_count", _I, [_VP]),
    ("f2dg_body_count", _I, [_VP, _I]),
    ("f2dg_shape_count", _I, [_VP, _I]),
    ("f2dg_awake_body_count", _L, [_VP]),
    ("f2dg_set_contact_capacity_factor", _I, [_VP, _F]),
    ("f2dg_set_stream", _I, [_VP, _VP]),
    ("f2dg_upload", _I, [_VP]),
    ("f2dg_device_bytes", _L, [_VP]),
    ("f2dg_step", _I, [_VP, _F, _I, _I, _I]),
    ("f2dg_step_worlds", _I, [_VP, _PI, _I, _F, _I, _I]),
    ("f2dg_step_each", _I, [_VP, _F, _I, _I]),
    ("f2dg_step_timed", _I, [_VP, _F, _I, _I, _PF]),
    ("f2dg_kernel_times", _I, [_VP, _F, _I, _I, _PF]),
    ("f2dg_phase_times", _I, [_VP, _F, _I, _I, _PF]),
    ("f2dg_synchronize", _I, [_VP]),
    ("f2dg_sync", _I, [_VP]),
    ("f2dg_get_bodies", _I, [_VP, _I, _PF]),
    ("f2dg_export_world", _L, [_VP, _I, _VP, _L]),
    ("f2dg_import_world", _I, [_VP, _I, _VP, _L]),
    ("f2dg_last_error", ctypes.c_char_p, [_VP]),
    ("f2dg_world_counters", _I, [_VP, _I, _PI]),
    ("f2dg_batch_stats", _I, [_VP, ctypes.POINTER(ctypes.c_long)]),
    ("f2dg_autotune", _I, [_VP, _F, _I, _I, _PF]),
    ("f2dg_set_online_tuning", _I, [_VP, _I]),
    ("f2dg_set_block", _I, [_VP, _I]),
    ("f2dg_get_block", _I, [_VP]),
    ("f2dg_set_class_block", _I, [_VP, _I, _I]),
    ("f2dg_set_class_solve_lds", _I, [_VP, _I, _I]),
    ("f2dg_error_bits", _I, [_VP]),
    ("f2dg_size_classes", _I, [_VP, _PI, _I]),
    ("f2dg_contact_event_counts", _I, [_VP, _I, _PI]),
    ("f2dg_contact_begin_events", _I, [_VP, _I, _PD]),
    ("f2dg_contact_end_events", _I, [_VP, _I, _PD]),
    ("f2dg_contact_hit_events", _I, [_VP, _I, _PD]),
    ("f2dg_body_events", _I, [_VP, _I, _PD]),
    ("f2d_create_world", _U, [_F, _F, _F, _F, _F, _F, _F, _F, _I, _I]),
    ("f2d_destroy_world", None, [_U]),
    ("f2d_world_is_valid", _I, [_U]),
    ("f2d_world_step", None, [_U, _F, _I]),
    ("f2d_world_set_gravity", None, [_U, _F, _F]),
    ("f2d_world_get_gravity", None, [_U, _PF]),
    ("f2d_world_enable_sleeping", None, [_U, _I]),
    ("f2d_world_is_sleeping_enabled", _I, [_U]),
    ("f2d_world_enable_continuous", None, [_U, _I]),
    ("f2d_world_is_continuous_enabled", _I, [_U]),
    ("f2d_create_body", None, [_U, _I, _F, _F, _F, _F, _F, _F, _F, _F, _F, _F, _F, ctypes.c_char_p, _I, _I, _I, _I,
                               _I, _I, _PI]),
    ("f2d_body_is_valid", _I, [ctypes.c_int32, _U]),
    ("f2d_body_get_position", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_body_get_rotation", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_body_get_linear_velocity", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_body_get_angular_velocity", _F, [ctypes.c_int32, _U]),
    ("f2d_body_set_linear_velocity", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_body_apply_force_to_center", None, [ctypes.c_int32, _U, _F, _F, _I]),
    ("f2d_body_apply_linear_impulse_to_center", None, [ctypes.c_int32, _U, _F, _F, _I]),
    ("f2d_body_get_world_center", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_body_get_local_center", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_body_is_awake", _I, [ctypes.c_int32, _U]),
    ("f2d_body_set_awake", None, [ctypes.c_int32, _U, _I]),
    ("f2d_body_is_sleep_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_body_enable_sleep", None, [ctypes.c_int32, _U, _I]),
    ("f2d_body_get_sleep_threshold", _F, [ctypes.c_int32, _U]),
    ("f2d_body_set_sleep_threshold", None, [ctypes.c_int32, _U, _F]),
    ("f2d_body_get_mass", _F, [ctypes.c_int32, _U]),
    ("f2d_body_get_rotational_inertia", _F, [ctypes.c_int32, _U]),
    ("f2d_body_get_type", _I, [ctypes.c_int32, _U]),
    ("f2d_create_circle_shape", None, [ctypes.c_int32, _U, _VP, _F, _F, _F, _PI]),
    ("f2d_create_capsule_shape", None, [ctypes.c_int32, _U, _VP, _F, _F, _F, _F, _F, _PI]),
    ("f2d_create_segment_shape", None, [ctypes.c_int32, _U, _VP, _F, _F, _F, _F, _PI]),
    ("f2d_create_box_shape", None, [ctypes.c_int32, _U, _VP, _F, _F, _F, _F, _F, _F, _F, _PI]),
    ("f2d_create_polygon_shape", _I, [ctypes.c_int32, _U, _VP, _PF, _I, _F, _PI]),
    ("f2d_create_revolute_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _F, _I, _F,
                                         _F, _I, _F, _F, _I, _F, _F, _F, _I, _PI]),
    ("f2d_world_get_contact_event_counts", None, [_U, _PI]),
    ("f2d_world_get_contact_begin_events", _I, [_U, _PD]),
    ("f2d_world_get_contact_end_events", _I, [_U, _PD]),
    ("f2d_world_get_contact_hit_events", _I, [_U, _PD]),
    ("f2d_world_get_body_event_count", _I, [_U]),
    ("f2d_world_get_body_events", _I, [_U, _PD]),
    ("f2d_body_set_transform", None, [ctypes.c_int32, _U, _F, _F, _F, _F]),
    ("f2d_body_set_angular_velocity", None, [ctypes.c_int32, _U, _F]),
    ("f2d_body_apply_force", None, [ctypes.c_int32, _U, _F, _F, _F, _F, _I]),
    ("f2d_body_apply_torque", None, [ctypes.c_int32, _U, _F, _I]),
    ("f2d_body_apply_linear_impulse", None, [ctypes.c_int32, _U, _F, _F, _F, _F, _I]),
    ("f2d_body_apply_angular_impulse", None, [ctypes.c_int32, _U, _F, _I]),
    ("f2d_body_set_mass_data", None, [ctypes.c_int32, _U, _F, _F, _F, _F]),
    ("f2d_body_apply_mass_from_shapes", None, [ctypes.c_int32, _U]),
    ("f2d_body_get_world_point", None, [ctypes.c_int32, _U, _F, _F, _PF]),
    ("f2d_body_get_local_point", None, [ctypes.c_int32, _U, _F, _F, _PF]),
    ("f2d_body_get_shape_count", _I, [ctypes.c_int32, _U]),
    ("f2d_body_get_shapes", _I, [ctypes.c_int32, _U, _PI, _I]),
    ("f2d_body_get_joint_count", _I, [ctypes.c_int32, _U]),
    ("f2d_body_get_joints", _I, [ctypes.c_int32, _U, _PI, _I]),
    ("f2d_body_get_name", ctypes.c_char_p, [ctypes.c_int32, _U]),
    ("f2d_body_set_name", None, [ctypes.c_int32, _U, ctypes.c_char_p]),
    ("f2d_body_is_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_body_get_gravity_scale", _F, [ctypes.c_int32, _U]),
    ("f2d_body_set_gravity_scale", None, [ctypes.c_int32, _U, _F]),
    ("f2d_body_get_linear_damping", _F, [ctypes.c_int32, _U]),
    ("f2d_body_set_linear_damping", None, [ctypes.c_int32, _U, _F]),
    ("f2d_body_get_angular_damping", _F, [ctypes.c_int32, _U]),
    ("f2d_body_set_angular_damping", None, [ctypes.c_int32, _U, _F]),
    ("f2d_body_is_bullet", _I, [ctypes.c_int32, _U]),
    ("f2d_body_set_bullet", None, [ctypes.c_int32, _U, _I]),
    ("f2d_body_is_fixed_rotation", _I, [ctypes.c_int32, _U]),
    ("f2d_body_set_fixed_rotation", None, [ctypes.c_int32, _U, _I]),
    ("f2d_shape_is_valid", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_get_body", None, [ctypes.c_int32, _U, _PI]),
    ("f2d_shape_get_type", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_is_sensor", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_get_density", _F, [ctypes.c_int32, _U]),
    ("f2d_shape_set_density", None, [ctypes.c_int32, _U, _F, _I]),
    ("f2d_shape_get_friction", _F, [ctypes.c_int32, _U]),
    ("f2d_shape_set_friction", None, [ctypes.c_int32, _U, _F]),
    ("f2d_shape_get_restitution", _F, [ctypes.c_int32, _U]),
    ("f2d_shape_set_restitution", None, [ctypes.c_int32, _U, _F]),
    ("f2d_shape_get_filter", None, [ctypes.c_int32, _U, _PI]),
    ("f2d_shape_are_contact_events_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_enable_contact_events", None, [ctypes.c_int32, _U, _I]),
    ("f2d_shape_are_hit_events_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_enable_hit_events", None, [ctypes.c_int32, _U, _I]),
    ("f2d_shape_are_pre_solve_events_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_shape_enable_pre_solve_events", None, [ctypes.c_int32, _U, _I]),
    ("f2d_shape_test_point", _I, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_shape_get_circle", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_shape_get_capsule", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_shape_get_segment", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_shape_get_polygon", _I, [ctypes.c_int32, _U, _PF]),
    ("f2d_shape_get_aabb", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_joint_is_valid", _I, [ctypes.c_int32, _U]),
    ("f2d_joint_get_type", _I, [ctypes.c_int32, _U]),
    ("f2d_joint_get_body_a", None, [ctypes.c_int32, _U, _PI]),
    ("f2d_joint_get_body_b", None, [ctypes.c_int32, _U, _PI]),
    ("f2d_joint_get_local_anchor_a", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_joint_get_local_anchor_b", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_joint_get_collide_connected", _I, [ctypes.c_int32, _U]),
    ("f2d_joint_wake_bodies", None, [ctypes.c_int32, _U]),
    ("f2d_joint_get_constraint_force", None, [ctypes.c_int32, _U, _PF]),
    ("f2d_joint_get_constraint_torque", _F, [ctypes.c_int32, _U]),
    ("f2d_create_prismatic_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _F, _F, _F,
                                          _I, _F, _F, _I, _F, _F, _I, _F, _F, _I, _PI]),
    ("f2d_prismatic_joint_is_spring_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_enable_spring", None, [ctypes.c_int32, _U, _I]),
    ("f2d_prismatic_joint_get_spring_hertz", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_spring_hertz", None, [ctypes.c_int32, _U, _F]),
    ("f2d_prismatic_joint_get_spring_damping_ratio", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_spring_damping_ratio", None, [ctypes.c_int32, _U, _F]),
    ("f2d_prismatic_joint_get_target_translation", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_target_translation", None, [ctypes.c_int32, _U, _F]),
    ("f2d_prismatic_joint_is_limit_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_enable_limit", None, [ctypes.c_int32, _U, _I]),
    ("f2d_prismatic_joint_get_lower_limit", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_get_upper_limit", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_limits", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_prismatic_joint_is_motor_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_enable_motor", None, [ctypes.c_int32, _U, _I]),
    ("f2d_prismatic_joint_get_motor_speed", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_motor_speed", None, [ctypes.c_int32, _U, _F]),
    ("f2d_prismatic_joint_get_max_motor_force", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_set_max_motor_force", None, [ctypes.c_int32, _U, _F]),
    ("f2d_prismatic_joint_get_motor_force", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_get_translation", _F, [ctypes.c_int32, _U]),
    ("f2d_prismatic_joint_get_speed", _F, [ctypes.c_int32, _U]),
    ("f2d_create_weld_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _F, _F, _F, _F,
                                     _I, _PI]),
    ("f2d_create_wheel_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _F, _I, _F, _F,
                                      _I, _F, _F, _I, _F, _F, _I, _PI]),
] + [(f"f2d_{k}_joint_get_{n}", _F, [ctypes.c_int32, _U]) for k, n in
     [("weld", "linear_hertz"), ("weld", "angular_hertz"), ("weld", "linear_damping_ratio"),
      ("weld", "angular_damping_ratio"), ("wheel", "spring_hertz"), ("wheel", "spring_damping_ratio"),
      ("wheel", "lower_limit"), ("wheel", "upper_limit"), ("wheel", "motor_speed"), ("wheel", "max_motor_torque"),
      ("wheel", "motor_torque")]] + [
    (f"f2d_{k}_joint_set_{n}", None, [ctypes.c_int32, _U, _F]) for k, n in
    [("weld", "linear_hertz"), ("weld", "angular_hertz"), ("weld", "linear_damping_ratio"),
     ("weld", "angular_damping_ratio"), ("wheel", "spring_hertz"), ("wheel", "spring_damping_ratio"),
     ("wheel", "motor_speed"), ("wheel", "max_motor_torque")]] + [
    ("f2d_create_distance_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _I, _F,
                                         _F, _I, _F, _F, _I, _F, _F, _I, _PI]),
] + [(f"f2d_distance_joint_get_{n}", _F, [ctypes.c_int32, _U]) for n in
     ["length", "current_length", "spring_hertz", "spring_damping_ratio", "min_length", "max_length", "motor_speed",
      "max_motor_force", "motor_force"]] + [
    (f"f2d_distance_joint_set_{n}", None, [ctypes.c_int32, _U, _F]) for n in
    ["length", "spring_hertz", "spring_damping_ratio", "motor_speed", "max_motor_force"]] + [
    (f"f2d_distance_joint_{n}", _I, [ctypes.c_int32, _U]) for n in
    ["is_spring_enabled", "is_limit_enabled", "is_motor_enabled"]] + [
    (f"f2d_distance_joint_{n}", None, [ctypes.c_int32, _U, _I]) for n in
    ["enable_spring", "enable_limit", "enable_motor"]] + [
    ("f2d_distance_joint_set_length_range", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_create_motor_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _F, _I,
                                      ctypes.POINTER(ctypes.c_int32)]),
    ("f2d_create_mouse_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, _F, _F, _F, _F, _F, _I,
                                      ctypes.POINTER(ctypes.c_int32)]),
    ("f2d_destroy_joint", None, [ctypes.c_int32, _U]),
    ("f2d_joint_set_collide_connected", None, [ctypes.c_int32, _U, _I]),
    ("f2d_create_filter_joint", None, [_U, ctypes.c_int32, _U, ctypes.c_int32, _U, ctypes.POINTER(ctypes.c_int32)]),
    ("f2d_motor_joint_get_linear_offset", None, [ctypes.c_int32, _U, ctypes.POINTER(ctypes.c_float)]),
    ("f2d_motor_joint_set_linear_offset", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_mouse_joint_get_target", None, [ctypes.c_int32, _U, ctypes.POINTER(ctypes.c_float)]),
    ("f2d_mouse_joint_set_target", None, [ctypes.c_int32, _U, _F, _F]),
] + [(f"f2d_{t}_joint_get_{n}", _F, [ctypes.c_int32, _U]) for t, n in
     [("motor", "angular_offset"), ("motor", "max_force"), ("motor", "max_torque"), ("motor", "correction_factor"),
      ("mouse", "spring_hertz"), ("mouse", "spring_damping_ratio"), ("mouse", "max_force")]] + [
    (f"f2d_{t}_joint_set_{n}", None, [ctypes.c_int32, _U, _F]) for t, n in
    [("motor", "angular_offset"), ("motor", "max_force"), ("motor", "max_torque"), ("motor", "correction_factor"),
     ("mouse", "spring_hertz"), ("mouse", "spring_damping_ratio"), ("mouse", "max_force")]] + [
    (f"f2d_revolute_joint_get_{n}", _F, [ctypes.c_int32, _U]) for n in
    ["spring_hertz", "spring_damping_ratio", "target_angle", "angle", "lower_limit", "upper_limit", "motor_speed",
     "max_motor_torque", "motor_torque"]] + [
    (f"f2d_revolute_joint_set_{n}", None, [ctypes.c_int32, _U, _F]) for n in
    ["spring_hertz", "spring_damping_ratio", "target_angle", "motor_speed", "max_motor_torque"]] + [
    (f"f2d_revolute_joint_{n}", _I, [ctypes.c_int32, _U]) for n in
    ["is_spring_enabled", "is_limit_enabled", "is_motor_enabled"]] + [
    (f"f2d_revolute_joint_{n}", None, [ctypes.c_int32, _U, _I]) for n in
    ["enable_spring", "enable_limit", "enable_motor"]] + [
    ("f2d_revolute_joint_set_limits", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_wheel_joint_is_spring_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_wheel_joint_enable_spring", None, [ctypes.c_int32, _U, _I]),
    ("f2d_wheel_joint_is_limit_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_wheel_joint_enable_limit", None, [ctypes.c_int32, _U, _I]),
    ("f2d_wheel_joint_set_limits", None, [ctypes.c_int32, _U, _F, _F]),
    ("f2d_wheel_joint_is_motor_enabled", _I, [ctypes.c_int32, _U]),
    ("f2d_wheel_joint_enable_motor", None, [ctypes.c_int32, _U, _I]),
    ("f2d_world_batch", _VP, [_U]),
    ("f2d_set_query_callbacks", None, [_VP, _VP]),
    ("f2d_world_cast_ray_closest", _I, [_U, _F, _F, _F, _F, _U, _U, _U, _U, _PD]),
    ("f2d_world_cast_ray", None, [_U, _F, _F, _F, _F, _U, _U, _U, _U]),
    ("f2d_world_overlap_aabb", None, [_U, _F, _F, _F, _F, _U, _U, _U, _U]),
    ("f2d_world_explode", None, [_U, _U, _U, _F, _F, _F, _F, _F]),
    ("f2dg_query", _I, [_VP, _VP, _I, _VP, _PI, _I]),
    ("f2dg_query_device", _I, [_VP, _VP, _I, _VP, _VP, _I]),
    ("f2dg_query_device_static", _I, [_VP, _VP, _I, _VP, _VP, _I]),
]

# query kinds and records of f2dg_query (include/forge2d_amd.h)
Q_RAY_CLOSEST, Q_RAY_ALL, Q_OVERLAP_AABB, Q_EXPLODE = 1, 2, 3, 4
QUERY_DTYPE = np.dtype([("world", "<i4"), ("kind", "<i4"), ("cat_lo", "<u4"), ("cat_hi", "<u4"),
                        ("mask_lo", "<u4"), ("mask_hi", "<u4"), ("p", "<f4", (6,))])
HIT_DTYPE = np.dtype([("shape", "<i4"), ("px", "<f4"), ("py", "<f4"), ("nx", "<f4"), ("ny", "<f4"),
                      ("fraction", "<f4")])
CAST_RAY_FN = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_float)
OVERLAP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int32, ctypes.c_uint32)
ALL_BITS = 0xFFFFFFFF


def make_queries(rows) -> np.ndarray:
    """rows of (world, kind, params[, (cat_lo, cat_hi, mask_lo, mask_hi)]) -> f2dQuery records."""
    q = np.zeros(len(rows), QUERY_DTYPE)
    for i, r in enumerate(rows):
        world, kind, params = r[0], r[1], r[2]
        bits = r[3] if len(r) > 3 else (1, 0, ALL_BITS, ALL_BITS)
        q[i]["world"], q[i]["kind"] = world, kind
        q[i]["cat_lo"], q[i]["cat_hi"], q[i]["mask_lo"], q[i]["mask_hi"] = bits
        q[i]["p"][:len(params)] = params
    return q


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load the HIP library once; raise loudly when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m forge2d_amd.build` (hipcc, gfx950)")
    try:
        l = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - environment specific
        raise LibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
    for name, res, args in SIGNATURES:
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    _lib = l
    return l


class Batch:
    """A set of independent worlds on one GPU, built from a scene spec."""

    def __init__(self, spec: bytes | list, capacity_factor: float | None = None):
        if not isinstance(spec, (bytes, bytearray)):
            spec = scenes.to_bytes(spec)
        self._spec = bytes(spec)
        L = lib()
        buf = ctypes.create_string_buffer(self._spec, len(self._spec))
        self.h = L.f2dg_create_from_spec(ctypes.cast(buf, ctypes.c_void_p), len(self._spec))
        if not self.h:
            raise ValueError("scene spec rejected by the library")
        if capacity_factor is not None:
            L.f2dg_set_contact_capacity_factor(self.h, capacity_factor)
        self.world_count = L.f2dg_world_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().f2dg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            err = lib().f2dg_last_error(self.h)
            raise RuntimeError(f"{what} failed ({rc}): {err.decode() if err else ''}")

    def upload(self):
        self._check(lib().f2dg_upload(self.h), "upload")

    def set_stream(self, stream_handle: int):
        """Launch on this hipStream_t (an integer handle, e.g. torch.cuda.Stream().cuda_stream)."""
        self._check(lib().f2dg_set_stream(self.h, ctypes.c_void_p(stream_handle)), "set_stream")

    def shape_count(self, world: int) -> int:
        return lib().f2dg_shape_count(self.h, world)

    def body_count(self, world: int) -> int:
        return lib().f2dg_body_count(self.h, world)

    def step(self, dt=1.0 / 60.0, substeps=4, steps=1, phases=PH_ALL):
        self._check(lib().f2dg_step(self.h, dt, substeps, steps, phases), "step")

    def step_worlds(self, worlds, dt=1.0 / 60.0, substeps=4, steps=1):
        ids = np.ascontiguousarray(worlds, dtype=np.int32)
        self._check(lib().f2dg_step_worlds(self.h, ids.ctypes.data_as(_PI), len(ids), dt, substeps, steps),
                    "step_worlds")

    def step_each(self, dt=1.0 / 60.0, substeps=4, steps=1):
        self._check(lib().f2dg_step_each(self.h, dt, substeps, steps), "step_each")

    def step_timed(self, dt=1.0 / 60.0, substeps=4, steps=1) -> float:
        ms = ctypes.c_float(0.0)
        self._check(lib().f2dg_step_timed(self.h, dt, substeps, steps, ctypes.byref(ms)), "step_timed")
        return ms.value

    def kernel_times(self, dt=1.0 / 60.0, substeps=4, steps=1) -> dict:
        """Step `steps` times; return summed HIP-event ms of each kernel."""
        ms = (ctypes.c_float * 4)()
        self._check(lib().f2dg_kernel_times(self.h, dt, substeps, steps, ms), "kernel_times")
        return {"collideKernel": ms[0], "solveKernel": ms[1], "continuousKernel": ms[2], "treeKernel": ms[3]}

    def phase_times(self, dt=1.0 / 60.0, substeps=4, steps=1) -> dict:
        """Summed milliseconds per phase over `steps` steps run as separate launches."""
        ms = (ctypes.c_float * 4)()
        self._check(lib().f2dg_phase_times(self.h, dt, substeps, steps, ms), "phase_times")
        return {"pairs": ms[0], "collide_state": ms[1], "solve_finalize": ms[2], "continuous_sleep": ms[3]}

    def synchronize(self):
        self._check(lib().f2dg_synchronize(self.h), "synchronize")

    def sync(self) -> int:
        rc = lib().f2dg_sync(self.h)
        if rc < 0:
            raise RuntimeError("sync failed")
        return rc

    def error_bits(self) -> int:
        """OR of the worlds' device error bits, from the world headers alone."""
        rc = lib().f2dg_error_bits(self.h)
        if rc < 0:
            raise RuntimeError("error_bits failed")
        return rc

    def bodies(self, world: int) -> np.ndarray:
        n = lib().f2dg_body_count(self.h, world)
        out = np.zeros((n, 7), np.float32)
        lib().f2dg_get_bodies(self.h, world, out.ctypes.data_as(_PF))
        return out

    def export(self, world: int) -> bytes:
        n = lib().f2dg_export_world(self.h, world, None, 0)
        if n < 0:
            raise RuntimeError("export failed")
        buf = ctypes.create_string_buffer(n)
        lib().f2dg_export_world(self.h, world, buf, n)
        return buf.raw

    def snapshot(self, world: int) -> snapshot.Snapshot:
        return snapshot.parse_one(self.export(world))

    def import_state(self, world: int, snap_bytes: bytes):
        buf = ctypes.create_string_buffer(snap_bytes, len(snap_bytes))
        self._check(lib().f2dg_import_world(self.h, world, buf, len(snap_bytes)), "import")

    def counters(self, world: int) -> dict:
        out = (ctypes.c_int32 * 8)()
        lib().f2dg_world_counters(self.h, world, out)
        keys = ["bodies", "shapes", "contacts", "touching", "overflow", "active_colors", "step", "errors"]
        return dict(zip(keys, list(out)))

    def autotune(self, dt=1.0 / 60.0, substeps=4, steps=4):
        ms = (ctypes.c_float * 3)()
        blk = lib().f2dg_autotune(self.h, dt, substeps, steps, ms)
        if blk < 0:
            raise RuntimeError("autotune failed")
        return blk, {64: ms[0], 128: ms[1], 256: ms[2]}

    def set_online_tuning(self, period: int):
        """Re-pick every size class's lanes per world every `period` steps (0: off)."""
        self._check(lib().f2dg_set_online_tuning(self.h, period), "set_online_tuning")

    def set_block(self, block: int):
        self._check(lib().f2dg_set_block(self.h, block), "set_block")

    def set_class_block(self, cls: int, block: int):
        self._check(lib().f2dg_set_class_block(self.h, cls, block), "set_class_block")

    def set_class_solve_lds(self, cls: int, nbytes: int):
        self._check(lib().f2dg_set_class_solve_lds(self.h, cls, nbytes), "set_class_solve_lds")

    @property
    def block(self) -> int:
        return lib().f2dg_get_block(self.h)

    def contact_events(self, world: int) -> dict:
        """Begin (n x 13), end (n x 4) and hit (n x 9) contact events of the last step,
        float64 records in the flat shim's layouts."""
        L = lib()
        c = (ctypes.c_int32 * 3)()
        self._check(L.f2dg_contact_event_counts(self.h, world, c), "contact_event_counts")
        out = {}
        for key, fn, width, n in (("begin", L.f2dg_contact_begin_events, 13, c[0]),
                                  ("end", L.f2dg_contact_end_events, 4, c[1]),
                                  ("hit", L.f2dg_contact_hit_events, 9, c[2])):
            a = np.zeros((n, width), np.float64)
            if n:
                fn(self.h, world, a.ctypes.data_as(_PD))
            out[key] = a
        return out

    def body_events(self, world: int) -> np.ndarray:
        """Body move events of the last step (n x 7)."""
        L = lib()
        n = L.f2dg_body_events(self.h, world, None)
        a = np.zeros((max(n, 0), 7), np.float64)
        if n > 0:
            L.f2dg_body_events(self.h, world, a.ctypes.data_as(_PD))
        return a

    def size_classes(self) -> list:
        """Launch size classes: limit, worlds, max bodies/shapes, workgroup lanes, LDS bytes."""
        out = (ctypes.c_int32 * (7 * 8))()
        n = lib().f2dg_size_classes(self.h, out, 8)
        if n < 0:
            raise RuntimeError("size_classes failed")
        keys = ("limit", "worlds", "max_bodies", "max_shapes", "lanes", "collide_lds", "solve_lds")
        return [dict(zip(keys, out[7 * k:7 * k + 7])) for k in range(n)]

    def stats(self) -> dict:
        out = (ctypes.c_long * 10)()
        self._check(lib().f2dg_batch_stats(self.h, out), "batch_stats")
        keys = ["bodies", "shapes", "contacts", "points", "restitution_contacts", "restitution_points", "joints",
                "overflow", "tree_proxies", "live_contacts"]
        return dict(zip(keys, list(out)))

    def device_bytes(self) -> int:
        return lib().f2dg_device_bytes(self.h)

    def query(self, queries: np.ndarray, cap: int = 64) -> tuple[np.ndarray, np.ndarray]:
        """Run f2dQuery records (make_queries); returns (counts, hits[n, cap])."""
        queries = np.ascontiguousarray(queries, QUERY_DTYPE)
        n = len(queries)
        hits = np.zeros((max(n, 1), cap), HIT_DTYPE)
        counts = np.zeros(max(n, 1), np.int32)
        self._check(lib().f2dg_query(self.h, queries.ctypes.data, n, hits.ctypes.data,
                                     counts.ctypes.data_as(_PI), cap), "query")
        return counts[:n], hits[:n]

    def awake_bodies(self) -> int:
        return lib().f2dg_awake_body_count(self.h)
