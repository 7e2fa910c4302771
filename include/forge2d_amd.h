/*
 * forge2d_amd — MI355X-native Box2D v3.1 world step behind a plain C-ABI.
 *
 * Library: forge2d_amd/libforge2d_amd.so (HIP for gfx950).
 *
 * Two families of entry points, both scalars/pointers/sizes only:
 *
 *  f2d_*   The flat per-world ABI of the reference's own C shim
 *          (packages/forge2d/native/wasm/f2d_shim.c, "F/native/wasm/f2d_shim.c"
 *          below): same names, argument order and meaning, same id packing
 *          (world ids = index1 | generation << 16; body/shape/joint ids =
 *          (int32 index1, uint32 world0 | generation << 16)). The step path,
 *          its events, sleep and the world queries are implemented;
 *          World.step maps to f2d_world_step, which
 *          replaces b2World_Step (packages/forge2d/third_party/box2d/src/world.c:695-812)
 *          as bound by RawBox2DFfi.worldStep (lib/src/backend/raw_box2d_ffi.dart:103-104).
 *
 *  f2dg_*  The batch ABI: many independent worlds packed in one device batch,
 *          stepped by one kernel launch per step (one workgroup per world).
 *          It has no counterpart in the reference, which steps one world per
 *          call.
 *
 * Errors: functions returning int return 0 on success and a negative value on
 * failure (f2dg_sync returns the OR of per-world error bits: 1 contact capacity,
 * 2 pair capacity, 4 overflow capacity). f2d_world_step aborts on a device error,
 * as the reference asserts (include/box2d/base.h B2_ASSERT).
 */
#ifndef FORGE2D_AMD_H
#define FORGE2D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- flat per-world ABI (F/native/wasm/f2d_shim.c) ------------------------ */

/* Shape definition record, F/native/wasm/f2d_shim.c:412-431 */
typedef struct f2dShapeDef {
  float friction;
  float restitution;
  float rolling_resistance;
  float tangent_speed;
  int32_t user_material_id;
  uint32_t custom_color;
  float density;
  uint32_t category_lo;
  uint32_t category_hi;
  uint32_t mask_lo;
  uint32_t mask_hi;
  int32_t group_index;
  int32_t is_sensor;
  int32_t enable_sensor_events;
  int32_t enable_contact_events;
  int32_t enable_hit_events;
  int32_t enable_pre_solve_events;
  int32_t invoke_contact_creation;
  int32_t update_body_mass;
} f2dShapeDef;

/* F/native/wasm/f2d_shim.c:123-139 (b2CreateWorld) */
uint32_t f2d_create_world(float gravity_x, float gravity_y, float restitution_threshold, float hit_event_threshold,
                          float contact_hertz, float contact_damping_ratio, float max_contact_push_speed,
                          float maximum_linear_speed, int enable_sleep, int enable_continuous);
/* F/native/wasm/f2d_shim.c:141 (b2DestroyWorld) */
void f2d_destroy_world(uint32_t w);
/* F/native/wasm/f2d_shim.c:143-145 (b2World_IsValid) */
int f2d_world_is_valid(uint32_t w);
/* F/native/wasm/f2d_shim.c:147-149 (b2World_Step) — the hot path */
void f2d_world_step(uint32_t w, float time_step, int sub_steps);
/* F/native/wasm/f2d_shim.c:151-159 (b2World_SetGravity / GetGravity) */
void f2d_world_set_gravity(uint32_t w, float x, float y);
void f2d_world_get_gravity(uint32_t w, float* out);
/* F/native/wasm/f2d_shim.c:161-175 (b2World_EnableSleeping / EnableContinuous) */
void f2d_world_enable_sleeping(uint32_t w, int enabled);
int f2d_world_is_sleeping_enabled(uint32_t w);
void f2d_world_enable_continuous(uint32_t w, int enabled);
int f2d_world_is_continuous_enabled(uint32_t w);

/* F/native/wasm/f2d_shim.c:192-218 (b2CreateBody); out = [index1, world_and_generation].
 * is_awake = 0 with enable_sleep = 1 creates the body asleep in a solver set and island of
 * its own, its velocities dropped, as the reference does (S/body.c:176-324).
 * is_enabled = 0 (the reference's disabled set) is not supported: out stays {0, 0} and
 * f2dg_last_error / stderr say so. */
void f2d_create_body(uint32_t w, int type, float px, float py, float qc, float qs, float lvx, float lvy,
                     float angular_velocity, float linear_damping, float angular_damping, float gravity_scale,
                     float sleep_threshold, const char* name, int enable_sleep, int is_awake, int fixed_rotation,
                     int is_bullet, int is_enabled, int allow_fast_rotation, int32_t* out);
/* F/native/wasm/f2d_shim.c:224-226 (b2Body_IsValid) */
int f2d_body_is_valid(int32_t index1, uint32_t world_and_generation);
/* F/native/wasm/f2d_shim.c:228-238 (b2Body_GetPosition / GetRotation) */
void f2d_body_get_position(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_body_get_rotation(int32_t index1, uint32_t world_and_generation, float* out);
/* F/native/wasm/f2d_shim.c:245-255 (b2Body_GetLinearVelocity / SetLinearVelocity) */
void f2d_body_get_linear_velocity(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_body_set_linear_velocity(int32_t index1, uint32_t world_and_generation, float x, float y);
/* F/native/wasm/f2d_shim.c (F2D_BODY_GET_FLOAT b2Body_GetAngularVelocity) */
float f2d_body_get_angular_velocity(int32_t index1, uint32_t world_and_generation);
/* F/native/wasm/f2d_shim.c:263-266 (b2Body_ApplyForceToCenter) */
void f2d_body_apply_force_to_center(int32_t index1, uint32_t world_and_generation, float fx, float fy, int wake);
/* F/native/wasm/f2d_shim.c:280-284 (b2Body_ApplyLinearImpulseToCenter) */
void f2d_body_apply_linear_impulse_to_center(int32_t index1, uint32_t world_and_generation, float ix, float iy,
                                             int wake);
/* F/native/wasm/f2d_shim.c:291-303 (b2Body_GetLocalCenterOfMass / GetWorldCenterOfMass) */
void f2d_body_get_world_center(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_body_get_local_center(int32_t index1, uint32_t world_and_generation, float* out);
/* F/native/wasm/f2d_shim.c (F2D_BODY_GET_FLOAT b2Body_GetMass / GetRotationalInertia / GetType) */
/* F/native/wasm/f2d_shim.c:389-394 (b2Body_IsAwake / SetAwake / IsSleepEnabled / EnableSleep /
 * Get/SetSleepThreshold): waking and putting to sleep run the step's island code on the device */
int f2d_body_is_awake(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_awake(int32_t index1, uint32_t world_and_generation, int awake);
int f2d_body_is_sleep_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_body_enable_sleep(int32_t index1, uint32_t world_and_generation, int enable);
float f2d_body_get_sleep_threshold(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_sleep_threshold(int32_t index1, uint32_t world_and_generation, float threshold);
float f2d_body_get_mass(int32_t index1, uint32_t world_and_generation);
float f2d_body_get_rotational_inertia(int32_t index1, uint32_t world_and_generation);
int f2d_body_get_type(int32_t index1, uint32_t world_and_generation);
/* F/native/wasm/f2d_shim.c:240-243 (b2Body_SetTransform, S/body.c:688-735): the shapes'
 * boxes follow; a proxy whose fat box no longer holds its box moves in its tree */
void f2d_body_set_transform(int32_t index1, uint32_t world_and_generation, float px, float py, float qc, float qs);
/* F/native/wasm/f2d_shim.c:381 (b2Body_SetAngularVelocity, S/body.c:791-812) */
void f2d_body_set_angular_velocity(int32_t index1, uint32_t world_and_generation, float w);
/* F/native/wasm/f2d_shim.c:257-289 (b2Body_ApplyForce / ApplyTorque / ApplyLinearImpulse /
 * ApplyAngularImpulse, S/body.c:900-1020): wake on request, awake bodies only; a linear
 * impulse is followed by the world's speed limit (b2LimitVelocity) */
void f2d_body_apply_force(int32_t index1, uint32_t world_and_generation, float fx, float fy, float px, float py,
                          int wake);
void f2d_body_apply_torque(int32_t index1, uint32_t world_and_generation, float torque, int wake);
void f2d_body_apply_linear_impulse(int32_t index1, uint32_t world_and_generation, float ix, float iy, float px,
                                   float py, int wake);
void f2d_body_apply_angular_impulse(int32_t index1, uint32_t world_and_generation, float impulse, int wake);
/* F/native/wasm/f2d_shim.c:305-310 (b2Body_SetMassData, S/body.c:1357-1381) and
 * F2D_BODY_CALL b2Body_ApplyMassFromShapes (S/body.c:1393-1403) */
void f2d_body_set_mass_data(int32_t index1, uint32_t world_and_generation, float mass, float rotational_inertia,
                            float cx, float cy);
void f2d_body_apply_mass_from_shapes(int32_t index1, uint32_t world_and_generation);
/* F/native/wasm/f2d_shim.c:320-332 (b2Body_GetWorldPoint / GetLocalPoint) */
void f2d_body_get_world_point(int32_t index1, uint32_t world_and_generation, float x, float y, float* out);
void f2d_body_get_local_point(int32_t index1, uint32_t world_and_generation, float x, float y, float* out);
/* F/native/wasm/f2d_shim.c:334-365 (b2Body_GetShapeCount / GetShapes / GetJointCount /
 * GetJoints): out = [index1, world_and_generation] per id, list order */
int f2d_body_get_shape_count(int32_t index1, uint32_t world_and_generation);
int f2d_body_get_shapes(int32_t index1, uint32_t world_and_generation, int32_t* out, int capacity);
int f2d_body_get_joint_count(int32_t index1, uint32_t world_and_generation);
int f2d_body_get_joints(int32_t index1, uint32_t world_and_generation, int32_t* out, int capacity);
/* F/native/wasm/f2d_shim.c:312-318 (b2Body_GetName / SetName, 31 characters) */
const char* f2d_body_get_name(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_name(int32_t index1, uint32_t world_and_generation, const char* name);
/* F/native/wasm/f2d_shim.c:395-407 (F2D_BODY_* accessors: enabled, gravity scale, damping,
 * bullet, fixed rotation; S/body.c:1405-1762). Every body here is enabled. */
int f2d_body_is_enabled(int32_t index1, uint32_t world_and_generation);
float f2d_body_get_gravity_scale(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_gravity_scale(int32_t index1, uint32_t world_and_generation, float v);
float f2d_body_get_linear_damping(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_linear_damping(int32_t index1, uint32_t world_and_generation, float v);
float f2d_body_get_angular_damping(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_angular_damping(int32_t index1, uint32_t world_and_generation, float v);
int f2d_body_is_bullet(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_bullet(int32_t index1, uint32_t world_and_generation, int flag);
int f2d_body_is_fixed_rotation(int32_t index1, uint32_t world_and_generation);
void f2d_body_set_fixed_rotation(int32_t index1, uint32_t world_and_generation, int flag);

/* F/native/wasm/f2d_shim.c:459-466 (b2CreateCircleShape) */
void f2d_create_circle_shape(int32_t index1, uint32_t world_and_generation, const f2dShapeDef* def, float cx,
                             float cy, float radius, int32_t* out);
/* F/native/wasm/f2d_shim.c:468-476 (b2CreateCapsuleShape) */
void f2d_create_capsule_shape(int32_t index1, uint32_t world_and_generation, const f2dShapeDef* def, float c1x,
                              float c1y, float c2x, float c2y, float radius, int32_t* out);
/* F/native/wasm/f2d_shim.c:478-486 (b2CreateSegmentShape) */
void f2d_create_segment_shape(int32_t index1, uint32_t world_and_generation, const f2dShapeDef* def, float p1x,
                              float p1y, float p2x, float p2y, int32_t* out);
/* F/native/wasm/f2d_shim.c:488-500 (b2MakeOffsetRoundedBox + b2CreatePolygonShape) */
void f2d_create_box_shape(int32_t index1, uint32_t world_and_generation, const f2dShapeDef* def, float half_width,
                          float half_height, float cx, float cy, float qc, float qs, float radius, int32_t* out);
/* F/native/wasm/f2d_shim.c:502-515 (b2ComputeHull + b2MakePolygon + b2CreatePolygonShape); 0 = degenerate */
int f2d_create_polygon_shape(int32_t index1, uint32_t world_and_generation, const f2dShapeDef* def,
                             const float* points, int count, float radius, int32_t* out);
/* F/native/wasm/f2d_shim.c:522-617 (shape accessors, S/shape.c:979-1340): body, type (b2ShapeType),
 * sensor, density (set: mass update on request), friction, restitution, filter
 * [category lo, hi, mask lo, hi, group], event flags, point test, geometry (polygon: radius,
 * then vertex pairs; returns the count), the shape's box */
int f2d_shape_is_valid(int32_t index1, uint32_t world_and_generation);
void f2d_shape_get_body(int32_t index1, uint32_t world_and_generation, int32_t* out);
int f2d_shape_get_type(int32_t index1, uint32_t world_and_generation);
int f2d_shape_is_sensor(int32_t index1, uint32_t world_and_generation);
float f2d_shape_get_density(int32_t index1, uint32_t world_and_generation);
void f2d_shape_set_density(int32_t index1, uint32_t world_and_generation, float density, int update_body_mass);
float f2d_shape_get_friction(int32_t index1, uint32_t world_and_generation);
void f2d_shape_set_friction(int32_t index1, uint32_t world_and_generation, float v);
float f2d_shape_get_restitution(int32_t index1, uint32_t world_and_generation);
void f2d_shape_set_restitution(int32_t index1, uint32_t world_and_generation, float v);
void f2d_shape_get_filter(int32_t index1, uint32_t world_and_generation, int32_t* out);
int f2d_shape_are_contact_events_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_shape_enable_contact_events(int32_t index1, uint32_t world_and_generation, int flag);
int f2d_shape_are_hit_events_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_shape_enable_hit_events(int32_t index1, uint32_t world_and_generation, int flag);
int f2d_shape_are_pre_solve_events_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_shape_enable_pre_solve_events(int32_t index1, uint32_t world_and_generation, int flag);
int f2d_shape_test_point(int32_t index1, uint32_t world_and_generation, float x, float y);
void f2d_shape_get_circle(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_shape_get_capsule(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_shape_get_segment(int32_t index1, uint32_t world_and_generation, float* out);
int f2d_shape_get_polygon(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_shape_get_aabb(int32_t index1, uint32_t world_and_generation, float* out);

/* F/native/wasm/f2d_shim.c:890-893 (b2DestroyJoint, S/joint.c:702-819): the joint
 * leaves its bodies, island and colour, and its id returns to the pool (the next joint
 * created takes it, with a new generation, so old handles become invalid). A joint of a
 * sleeping set or between two static bodies is left alone with a message on stderr. */
void f2d_destroy_joint(int32_t index1, uint32_t world_and_generation);

/* F/native/wasm/f2d_shim.c:932-935 (b2Joint_SetCollideConnected, S/joint.c:979-1022):
 * turning it on buffers the shapes of the body with fewer shapes as moved; turning it
 * off between bodies that share a contact is refused with a message on stderr (the
 * reference destroys those contacts, which is not restated here). */
void f2d_joint_set_collide_connected(int32_t index1, uint32_t world_and_generation, int v);

/* F/native/wasm/f2d_shim.c:816-843 (b2CreateRevoluteJoint).
 * Every joint creation call with collide_connected = 0 (revolute, prismatic, weld,
 * wheel, distance, motor) between bodies that already share a contact is refused: out
 * receives a null id and a message goes to stderr, since the reference would destroy
 * that contact (S/joint.c:313-345) and that edit is not restated here. Mouse and filter
 * joints destroy no contacts and are always made. */
void f2d_create_revolute_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                               uint32_t b_world_and_generation, float anchor_ax, float anchor_ay, float anchor_bx,
                               float anchor_by, float reference_angle, float target_angle, int enable_spring,
                               float hertz, float damping_ratio, int enable_limit, float lower_angle,
                               float upper_angle, int enable_motor, float max_motor_torque, float motor_speed,
                               float draw_size, int collide_connected, int32_t* out);

/* F/native/wasm/f2d_shim.c:787-814 (b2CreatePrismaticJoint, S/joint.c:559-610): the axis
 * is given in body A's frame and normalized */
void f2d_create_prismatic_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                                uint32_t b_world_and_generation, float anchor_ax, float anchor_ay, float anchor_bx,
                                float anchor_by, float axis_x, float axis_y, float reference_angle,
                                float target_translation, int enable_spring, float hertz, float damping_ratio,
                                int enable_limit, float lower_translation, float upper_translation, int enable_motor,
                                float max_motor_force, float motor_speed, int collide_connected, int32_t* out);

/* F/native/wasm/f2d_shim.c:1054-1094 (prismatic joint parameters, S/prismatic_joint.c:12-209);
 * a call on a joint of another type is ignored and getters return 0 */
int f2d_prismatic_joint_is_spring_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_enable_spring(int32_t index1, uint32_t world_and_generation, int v);
float f2d_prismatic_joint_get_spring_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_spring_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_prismatic_joint_get_spring_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_spring_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
float f2d_prismatic_joint_get_target_translation(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_target_translation(int32_t index1, uint32_t world_and_generation, float v);
int f2d_prismatic_joint_is_limit_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_enable_limit(int32_t index1, uint32_t world_and_generation, int v);
float f2d_prismatic_joint_get_lower_limit(int32_t index1, uint32_t world_and_generation);
float f2d_prismatic_joint_get_upper_limit(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_limits(int32_t index1, uint32_t world_and_generation, float lower, float upper);
int f2d_prismatic_joint_is_motor_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_enable_motor(int32_t index1, uint32_t world_and_generation, int v);
float f2d_prismatic_joint_get_motor_speed(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_motor_speed(int32_t index1, uint32_t world_and_generation, float v);
float f2d_prismatic_joint_get_max_motor_force(int32_t index1, uint32_t world_and_generation);
void f2d_prismatic_joint_set_max_motor_force(int32_t index1, uint32_t world_and_generation, float v);
float f2d_prismatic_joint_get_motor_force(int32_t index1, uint32_t world_and_generation);
float f2d_prismatic_joint_get_translation(int32_t index1, uint32_t world_and_generation);
float f2d_prismatic_joint_get_speed(int32_t index1, uint32_t world_and_generation);

/* F/native/wasm/f2d_shim.c:1097-1136 (revolute joint parameters, S/revolute_joint.c:20-176);
 * a call on a joint of another type is ignored and getters return 0 */
int f2d_revolute_joint_is_spring_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_enable_spring(int32_t index1, uint32_t world_and_generation, int v);
float f2d_revolute_joint_get_spring_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_spring_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_revolute_joint_get_spring_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_spring_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
float f2d_revolute_joint_get_target_angle(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_target_angle(int32_t index1, uint32_t world_and_generation, float v);
float f2d_revolute_joint_get_angle(int32_t index1, uint32_t world_and_generation);
int f2d_revolute_joint_is_limit_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_enable_limit(int32_t index1, uint32_t world_and_generation, int v);
float f2d_revolute_joint_get_lower_limit(int32_t index1, uint32_t world_and_generation);
float f2d_revolute_joint_get_upper_limit(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_limits(int32_t index1, uint32_t world_and_generation, float lower, float upper);
int f2d_revolute_joint_is_motor_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_enable_motor(int32_t index1, uint32_t world_and_generation, int v);
float f2d_revolute_joint_get_motor_speed(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_motor_speed(int32_t index1, uint32_t world_and_generation, float v);
float f2d_revolute_joint_get_max_motor_torque(int32_t index1, uint32_t world_and_generation);
void f2d_revolute_joint_set_max_motor_torque(int32_t index1, uint32_t world_and_generation, float v);
float f2d_revolute_joint_get_motor_torque(int32_t index1, uint32_t world_and_generation);

/* F/native/wasm/f2d_shim.c:716-740 (b2CreateDistanceJoint, S/joint.c:353-405) */
void f2d_create_distance_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                               uint32_t b_world_and_generation, float anchor_ax, float anchor_ay, float anchor_bx,
                               float anchor_by, float length, int enable_spring, float hertz, float damping_ratio,
                               int enable_limit, float min_length, float max_length, int enable_motor,
                               float max_motor_force, float motor_speed, int collide_connected, int32_t* out);

/* F/native/wasm/f2d_shim.c:751-768 (b2CreateMotorJoint, S/joint.c:406-442) */
void f2d_create_motor_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                            uint32_t b_world_and_generation, float offset_x, float offset_y, float angular_offset,
                            float max_force, float max_torque, float correction_factor, int collide_connected,
                            int32_t* out);
/* F/native/wasm/f2d_shim.c:770-785 (b2CreateMouseJoint, S/joint.c:444-478) */
void f2d_create_mouse_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                            uint32_t b_world_and_generation, float target_x, float target_y, float hertz,
                            float damping_ratio, float max_force, int collide_connected, int32_t* out);
/* F/native/wasm/f2d_shim.c:742-749 (b2CreateFilterJoint, S/joint.c:480-505) */
void f2d_create_filter_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                             uint32_t b_world_and_generation, int32_t* out);

/* F/native/wasm/f2d_shim.c:1006-1050 (motor and mouse joint parameters,
 * S/motor_joint.c:14-71, S/mouse_joint.c:14-63); a call on a joint of another type is
 * ignored and getters return 0 */
void f2d_motor_joint_get_linear_offset(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_motor_joint_set_linear_offset(int32_t index1, uint32_t world_and_generation, float x, float y);
float f2d_motor_joint_get_angular_offset(int32_t index1, uint32_t world_and_generation);
void f2d_motor_joint_set_angular_offset(int32_t index1, uint32_t world_and_generation, float v);
float f2d_motor_joint_get_max_force(int32_t index1, uint32_t world_and_generation);
void f2d_motor_joint_set_max_force(int32_t index1, uint32_t world_and_generation, float v);
float f2d_motor_joint_get_max_torque(int32_t index1, uint32_t world_and_generation);
void f2d_motor_joint_set_max_torque(int32_t index1, uint32_t world_and_generation, float v);
float f2d_motor_joint_get_correction_factor(int32_t index1, uint32_t world_and_generation);
void f2d_motor_joint_set_correction_factor(int32_t index1, uint32_t world_and_generation, float v);
void f2d_mouse_joint_get_target(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_mouse_joint_set_target(int32_t index1, uint32_t world_and_generation, float x, float y);
float f2d_mouse_joint_get_spring_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_mouse_joint_set_spring_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_mouse_joint_get_spring_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_mouse_joint_set_spring_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
float f2d_mouse_joint_get_max_force(int32_t index1, uint32_t world_and_generation);
void f2d_mouse_joint_set_max_force(int32_t index1, uint32_t world_and_generation, float v);

/* F/native/wasm/f2d_shim.c:965-1004 (distance joint parameters, S/distance_joint.c:19-183);
 * a call on a joint of another type is ignored and getters return 0 */
float f2d_distance_joint_get_length(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_length(int32_t index1, uint32_t world_and_generation, float v);
float f2d_distance_joint_get_current_length(int32_t index1, uint32_t world_and_generation);
int f2d_distance_joint_is_spring_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_enable_spring(int32_t index1, uint32_t world_and_generation, int v);
float f2d_distance_joint_get_spring_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_spring_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_distance_joint_get_spring_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_spring_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
int f2d_distance_joint_is_limit_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_enable_limit(int32_t index1, uint32_t world_and_generation, int v);
float f2d_distance_joint_get_min_length(int32_t index1, uint32_t world_and_generation);
float f2d_distance_joint_get_max_length(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_length_range(int32_t index1, uint32_t world_and_generation, float min_length,
                                         float max_length);
int f2d_distance_joint_is_motor_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_enable_motor(int32_t index1, uint32_t world_and_generation, int v);
float f2d_distance_joint_get_motor_speed(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_motor_speed(int32_t index1, uint32_t world_and_generation, float v);
float f2d_distance_joint_get_max_motor_force(int32_t index1, uint32_t world_and_generation);
void f2d_distance_joint_set_max_motor_force(int32_t index1, uint32_t world_and_generation, float v);
float f2d_distance_joint_get_motor_force(int32_t index1, uint32_t world_and_generation);

/* F/native/wasm/f2d_shim.c:844-862 (b2CreateWeldJoint, S/joint.c:612-649) */
void f2d_create_weld_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                           uint32_t b_world_and_generation, float anchor_ax, float anchor_ay, float anchor_bx,
                           float anchor_by, float reference_angle, float linear_hertz, float angular_hertz,
                           float linear_damping_ratio, float angular_damping_ratio, int collide_connected,
                           int32_t* out);

/* F/native/wasm/f2d_shim.c:864-888 (b2CreateWheelJoint, S/joint.c:651-700): the axis is
 * given in body A's frame and normalized */
void f2d_create_wheel_joint(uint32_t w, int32_t a_index1, uint32_t a_world_and_generation, int32_t b_index1,
                            uint32_t b_world_and_generation, float anchor_ax, float anchor_ay, float anchor_bx,
                            float anchor_by, float axis_x, float axis_y, int enable_spring, float hertz,
                            float damping_ratio, int enable_limit, float lower_translation, float upper_translation,
                            int enable_motor, float max_motor_torque, float motor_speed, int collide_connected,
                            int32_t* out);

/* F/native/wasm/f2d_shim.c:1138-1185 (weld and wheel joint parameters, S/weld_joint.c:14-64,
 * S/wheel_joint.c:12-145); a call on a joint of another type is ignored and getters return 0 */
float f2d_weld_joint_get_linear_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_weld_joint_set_linear_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_weld_joint_get_angular_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_weld_joint_set_angular_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_weld_joint_get_linear_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_weld_joint_set_linear_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
float f2d_weld_joint_get_angular_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_weld_joint_set_angular_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
int f2d_wheel_joint_is_spring_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_enable_spring(int32_t index1, uint32_t world_and_generation, int v);
float f2d_wheel_joint_get_spring_hertz(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_set_spring_hertz(int32_t index1, uint32_t world_and_generation, float v);
float f2d_wheel_joint_get_spring_damping_ratio(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_set_spring_damping_ratio(int32_t index1, uint32_t world_and_generation, float v);
int f2d_wheel_joint_is_limit_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_enable_limit(int32_t index1, uint32_t world_and_generation, int v);
float f2d_wheel_joint_get_lower_limit(int32_t index1, uint32_t world_and_generation);
float f2d_wheel_joint_get_upper_limit(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_set_limits(int32_t index1, uint32_t world_and_generation, float lower, float upper);
int f2d_wheel_joint_is_motor_enabled(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_enable_motor(int32_t index1, uint32_t world_and_generation, int v);
float f2d_wheel_joint_get_motor_speed(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_set_motor_speed(int32_t index1, uint32_t world_and_generation, float v);
float f2d_wheel_joint_get_max_motor_torque(int32_t index1, uint32_t world_and_generation);
void f2d_wheel_joint_set_max_motor_torque(int32_t index1, uint32_t world_and_generation, float v);
float f2d_wheel_joint_get_motor_torque(int32_t index1, uint32_t world_and_generation);

/* F/native/wasm/f2d_shim.c:892-950 (joint accessors, S/joint.c): type (0 b2_distanceJoint,
 * 4 b2_prismaticJoint, 5 b2_revoluteJoint, 6 b2_weldJoint, 7 b2_wheelJoint), bodies, local
 * anchors, collideConnected, waking both bodies, and the last step's constraint force and
 * torque (S/revolute_joint.c:176-187, S/prismatic_joint.c:211-232, S/weld_joint.c:66-75,
 * S/wheel_joint.c:147-165, S/distance_joint.c:185-198) */
int f2d_joint_is_valid(int32_t index1, uint32_t world_and_generation);
int f2d_joint_get_type(int32_t index1, uint32_t world_and_generation);
void f2d_joint_get_body_a(int32_t index1, uint32_t world_and_generation, int32_t* out);
void f2d_joint_get_body_b(int32_t index1, uint32_t world_and_generation, int32_t* out);
void f2d_joint_get_local_anchor_a(int32_t index1, uint32_t world_and_generation, float* out);
void f2d_joint_get_local_anchor_b(int32_t index1, uint32_t world_and_generation, float* out);
int f2d_joint_get_collide_connected(int32_t index1, uint32_t world_and_generation);
void f2d_joint_wake_bodies(int32_t index1, uint32_t world_and_generation);
void f2d_joint_get_constraint_force(int32_t index1, uint32_t world_and_generation, float* out);
float f2d_joint_get_constraint_torque(int32_t index1, uint32_t world_and_generation);

/* Events of the last step, F/native/wasm/f2d_shim.c:1186-1300 (counts first, then the
 * fill functions write float64 records and return the record count):
 *   begin 13 [aI1, aWg, bI1, bWg, nx, ny, pointCount, p0x, p0y, sep0, p1x, p1y, sep1]
 *   end    4 [aI1, aWg, bI1, bWg]
 *   hit    9 [aI1, aWg, bI1, bWg, px, py, nx, ny, approachSpeed]
 *   body   7 [index1, world_and_generation, x, y, qc, qs, fellAsleep] */
void f2d_world_get_contact_event_counts(uint32_t w, int32_t* out);
int f2d_world_get_contact_begin_events(uint32_t w, double* out);
int f2d_world_get_contact_end_events(uint32_t w, double* out);
int f2d_world_get_contact_hit_events(uint32_t w, double* out);
int f2d_world_get_body_event_count(uint32_t w);
int f2d_world_get_body_events(uint32_t w, double* out);

/* Queries, F/native/wasm/f2d_shim.c:1308-1367 (rays, AABB overlap) and :177-189
 * (explosion). A device kernel walks the world's trees as the reference's world queries
 * do; hit shapes, points, normals and fractions are the reference's bit for bit and
 * come in the reference's walk order. The reference calls the host imports
 * f2d_host_cast_ray / f2d_host_overlap (f2d_shim.c:27-31) during its tree walks; here
 * the host registers them with f2d_set_query_callbacks and they are called after the
 * device pass, in the walk order (static, kinematic, then dynamic tree), with the
 * reference's clipping rules (a return of 0 ends the cast, -1 skips the hit, a fraction
 * clips the ray for the hits after it; a false overlap return ends that tree).
 * f2d_world_cast_ray_closest writes [index1, world_and_generation, px, py, nx, ny,
 * fraction] and returns 1 on a hit. f2d_world_explode wakes and pushes the dynamic
 * bodies in reach in the dynamic tree's depth-first order. */
typedef float (*f2dHostCastRayFn)(int32_t shape_index1, uint32_t shape_wg, float px, float py, float nx, float ny,
                                  float fraction);
typedef int (*f2dHostOverlapFn)(int32_t shape_index1, uint32_t shape_wg);
void f2d_set_query_callbacks(f2dHostCastRayFn cast_ray, f2dHostOverlapFn overlap);
int f2d_world_cast_ray_closest(uint32_t w, float ox, float oy, float tx, float ty, uint32_t category_lo,
                               uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi, double* out);
void f2d_world_cast_ray(uint32_t w, float ox, float oy, float tx, float ty, uint32_t category_lo,
                        uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi);
void f2d_world_overlap_aabb(uint32_t w, float lx, float ly, float ux, float uy, uint32_t category_lo,
                            uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi);
void f2d_world_explode(uint32_t w, uint32_t mask_lo, uint32_t mask_hi, float x, float y, float radius,
                       float falloff, float impulse_per_length);

/* The batch handle behind a single f2d world (no counterpart in the reference):
 * lets the f2dg_* export/counter calls read a world created through f2d_*. */
void* f2d_world_batch(uint32_t w);

/* ---- batch ABI (no counterpart in the reference) --------------------------- */

/* Build worlds from a scene spec buffer (forge2d_amd/scenes.py layout). NULL on error. */
void* f2dg_create_from_spec(const void* spec, long bytes);
void f2dg_destroy(void* batch);
int f2dg_world_count(void* batch);
int f2dg_body_count(void* batch, int world);
int f2dg_shape_count(void* batch, int world);
long f2dg_awake_body_count(void* batch);
/* contact slots per shape reserved on the device (default 6, before upload) */
int f2dg_set_contact_capacity_factor(void* batch, float factor);
/* hipStream_t to launch on (NULL = default stream) */
int f2dg_set_stream(void* batch, void* stream);
/* lay out and copy the batch to the device (done implicitly by the first step) */
int f2dg_upload(void* batch);
long f2dg_device_bytes(void* batch);
/* step every world `steps` times; phases bitmask: 1 pairs, 2 collide, 4 contact state,
 * 8 solve, 16 finalize (0 = all) */
int f2dg_step(void* batch, float dt, int sub_steps, int steps, int phases);
/* step only the listed worlds (continuous-batching serving); the worlds of each
 * size class launch with that class's configuration on the class's stream */
int f2dg_step_worlds(void* batch, const int32_t* world_ids, int n, float dt, int sub_steps, int steps);
/* the one-world-at-a-time GPU path: a launch per world per step */
int f2dg_step_each(void* batch, float dt, int sub_steps, int steps);
/* step and time with HIP events on the batch stream; *ms_out = total milliseconds */
int f2dg_step_timed(void* batch, float dt, int sub_steps, int steps, float* ms_out);
/* step and time each kernel: HIP events bracket the collide, tree, solve and
 * continuous-collision launches of every step; summed milliseconds in ms[0] collide,
 * ms[1] solve, ms[2] continuous collision, ms[3] tree */
int f2dg_kernel_times(void* batch, float dt, int sub_steps, int steps, float* ms);
/* step with the phases as separate launches and time each: ms[0] pairs, ms[1] narrow
 * phase + contact state, ms[2] solver + finalize, ms[3] continuous + island sleep */
int f2dg_phase_times(void* batch, float dt, int sub_steps, int steps, float* ms);
int f2dg_synchronize(void* batch);
/* copy device state to the host mirror; returns OR of world error bits, <0 on failure */
int f2dg_sync(void* batch);
/* OR of the world error bits read from the device world headers alone (no mirror copy) */
int f2dg_error_bits(void* batch);
/* per body: px, py, qc, qs, vx, vy, w */
int f2dg_get_bodies(void* batch, int world, float* out);
/* world snapshot (forge2d_amd/snapshot.py layout); returns the size, writes when cap suffices */
long f2dg_export_world(void* batch, int world, void* out, long cap);
/* overlay a snapshot onto a world built from the same scene spec */
int f2dg_import_world(void* batch, int world, const void* snapshot, long bytes);
const char* f2dg_last_error(void* batch);
/* autotuner: per size class, times `steps` steps with 64, 128 and 256 lanes per world (and
 * 512 and 1024 for classes of more than 256 shapes) and keeps the fastest (ms_out[3]
 * optional: whole-batch ms per 64/128/256 candidate), then times 4 solve
 * LDS sizes (room for 0, 0.5, 1 and 2 constraint lanes per body) and keeps the fastest;
 * classes with 3 lane candidates run 2 x steps untimed steps more, so every world advances by
 * 9 x steps trial steps; returns the largest class's lanes */
int f2dg_autotune(void* batch, float dt, int sub_steps, int steps, float* ms_out);
/* online lane tuning of f2dg_step: every `period` steps (0: off), a round of three ordinary
 * steps runs each size class at 64, 128 and 256 lanes per world (256, 512 and 1024 for
 * classes of more than 512 shapes), the classes one after another, timed with HIP events,
 * and each class keeps its fastest; as worlds settle and gather contacts the best size
 * changes, and the lanes never change results */
int f2dg_set_online_tuning(void* batch, int period);
/* worlds are launched in size classes by max(bodies, shapes) <= 64, 128, ..., 1024, larger;
 * per class: [limit, worlds, max bodies, max shapes, workgroup lanes, collide LDS bytes,
 * solve LDS bytes] into out (7 ints each, up to cap classes); returns the class count */
int f2dg_size_classes(void* batch, int32_t* out, int cap);
/* force a workgroup size (64, 128, 256, 512 or 1024) for every size class, before or after the
 * batch is laid out / read the current one */
int f2dg_set_block(void* batch, int block);
int f2dg_get_block(void* batch);
/* force the workgroup size of size class `cls` (index into f2dg_size_classes) */
int f2dg_set_class_block(void* batch, int cls, int block);
/* force the solve launch's dynamic LDS bytes of size class `cls` (body terms of its
 * largest world + constraint lanes; worlds whose lanes do not fit keep them in HBM) */
int f2dg_set_class_solve_lds(void* batch, int cls, int bytes);
/* aggregate counts after f2dg_sync (10 longs): [awake bodies, shapes, touching contacts,
 * touching points, restitution contacts, restitution points, joints, overflow contacts,
 * proxies of the kinematic and dynamic trees, live contacts (touching or not)] */
int f2dg_batch_stats(void* batch, long* out);
/* events of one world's last step, in the f2d_world_get_*_events layouts above (shape and
 * body world_and_generation fields are 0); counts out = [begin, end, hit] */
int f2dg_contact_event_counts(void* batch, int world, int32_t* out);
int f2dg_contact_begin_events(void* batch, int world, double* out);
int f2dg_contact_end_events(void* batch, int world, double* out);
int f2dg_contact_hit_events(void* batch, int world, double* out);
int f2dg_body_events(void* batch, int world, double* out);
/* [bodies, shapes, live contacts, touching, overflow, active colours, step, error bits] */
int f2dg_world_counters(void* batch, int world, int32_t* out);

/* World queries over the worlds of a batch, one lane per query walking the world's
 * trees as the reference's world queries do (the reference queries one world per
 * call); results come in the reference's order. kind 1: closest ray hit
 * (b2World_CastRayClosest, S/world.c:2277-2308; p = origin x/y, translation x/y);
 * 2: every ray hit (b2World_CastRay with a callback returning 1, :2222-2254), in the
 * order of b2DynamicTree_RayCast's walks (static, kinematic, dynamic tree; nearer child
 * first); 3: every shape whose fat AABB overlaps the box p = lower x/y, upper x/y
 * (b2World_OverlapAABB, :2071-2105), in b2DynamicTree_Query's depth-first order;
 * 4: explosion (b2World_Explode, :2640-2747; p = position x/y, radius, falloff,
 * impulse per length; mask_lo/hi = the explosion's mask), the dynamic tree's shapes in
 * depth-first order. A count of -2 reports a tree deeper than the 64-entry walk
 * stack (a height above 63). */
typedef struct f2dQuery {
  int32_t world;
  int32_t kind;
  uint32_t category_lo, category_hi, mask_lo, mask_hi;
  float p[6];
} f2dQuery;
typedef struct f2dQueryHit {
  int32_t shape; /* world-local shape id (index1 - 1) */
  float px, py, nx, ny, fraction;
} f2dQueryHit;
/* Host arrays. Runs rays and overlaps in one launch, then the explosions in query
 * order (one launch per round, each world at most once per round). counts[i] = the
 * number of results of query i (0 for explosions, -1 for a world outside the batch);
 * the first min(count, cap) are written to hits[i * cap ...]. cap >= 1. */
int f2dg_query(void* batch, const f2dQuery* queries, int n, f2dQueryHit* hits, int32_t* counts, int cap);
/* Device arrays (queries, hits, counts), kinds 1-3 only, on the batch's stream with no
 * synchronization: the trees' pending leaf enlargements applied, the queries grouped by
 * world (a counting sort into a library buffer of n + worlds + 1 ints, grown on
 * demand), then one query per lane. Results do not depend on the grouping. */
int f2dg_query_device(void* batch, const void* queries, int n, void* hits, void* counts, int cap);
/* No counterpart in the reference: f2dg_query_device for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The grouping scratch is a static device array of 1,048,576 ints (4 MiB per GPU, held
 * until the process ends) instead of a library buffer; nothing is allocated. The
 * batch must be uploaded with no host edits pending, and n + worlds + 1 must be at
 * most 1,048,576; otherwise -1 and f2dg_last_error says why. Calls on one GPU share
 * the array: calls whose launches overlap on one GPU (other streams or threads) race. */
int f2dg_query_device_static(void* batch, const void* queries, int n, void* hits, void* counts, int cap);

#ifdef __cplusplus
}
#endif

#endif /* FORGE2D_AMD_H */
