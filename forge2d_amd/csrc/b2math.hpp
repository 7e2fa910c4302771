// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Scalar 2D math with the reference's exact operation order.
//
// Every function here evaluates the same IEEE-754 single-precision operations in
// the same order as its counterpart in I/math_functions.h (cited per function),
// so that a kernel compiled with -ffp-contract=off and IEEE division/sqrt gives
// the same bits as the reference's x86-64 SSE2 build. Min/max are written as
// the reference's ternaries (I/math_functions.h:126-135), never fminf/fmaxf,
// which differ on signed zeros and NaN.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define F2D_HD __host__ __device__ __forceinline__
// serial, rarely taken bookkeeping (island graph): kept out of line so its registers
// and stack do not weigh on the lanes of the kernels that call it
#define F2D_HD_COLD __host__ __device__ __noinline__ inline
#else
#define F2D_HD inline
#define F2D_HD_COLD inline
#endif

namespace f2d {

constexpr float kPi = 3.14159265359f;                 // B2_PI, I/math_functions.h:79
constexpr float kFltEpsilon = 1.1920928955078125e-7f;  // FLT_EPSILON
constexpr float kFltMax = 3.402823466e+38f;            // FLT_MAX
constexpr float kLinearSlop = 0.005f;                  // B2_LINEAR_SLOP, S/constants.h
constexpr float kSpeculativeDistance = 4.0f * kLinearSlop;  // B2_SPECULATIVE_DISTANCE
constexpr float kAabbMargin = 0.05f;                   // B2_AABB_MARGIN
constexpr float kHuge = 100000.0f;                     // B2_HUGE
constexpr float kMaxRotation = 0.25f * kPi;            // B2_MAX_ROTATION
constexpr float kTimeToSleep = 0.5f;                   // B2_TIME_TO_SLEEP
constexpr int kMaxPolygonVertices = 8;                 // B2_MAX_POLYGON_VERTICES
constexpr int kGraphColorCount = 12;                   // B2_GRAPH_COLOR_COUNT
constexpr int kOverflowIndex = kGraphColorCount - 1;   // B2_OVERFLOW_INDEX

struct Vec2 {
  float x, y;
};
struct Rot {
  float c, s;
};
struct Xf {
  Vec2 p;
  Rot q;
};
struct AABB {
  Vec2 lo, hi;
};

F2D_HD float minf(float a, float b) { return a < b ? a : b; }  // b2MinFloat :126-129
F2D_HD float maxf(float a, float b) { return a > b ? a : b; }  // b2MaxFloat :132-135
F2D_HD float absf(float a) { return a < 0 ? -a : a; }          // b2AbsFloat :138-141
F2D_HD float clampf(float a, float lo, float hi) { return a < lo ? lo : (a > hi ? hi : a); }
F2D_HD int mini(int a, int b) { return a < b ? a : b; }
F2D_HD int maxi(int a, int b) { return a > b ? a : b; }

F2D_HD float dot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
F2D_HD float cross(Vec2 a, Vec2 b) { return a.x * b.y - a.y * b.x; }
F2D_HD Vec2 crossVS(Vec2 v, float s) { return {s * v.y, -s * v.x}; }
F2D_HD Vec2 crossSV(float s, Vec2 v) { return {-s * v.y, s * v.x}; }
F2D_HD Vec2 leftPerp(Vec2 v) { return {-v.y, v.x}; }
F2D_HD Vec2 rightPerp(Vec2 v) { return {v.y, -v.x}; }
F2D_HD Vec2 add(Vec2 a, Vec2 b) { return {a.x + b.x, a.y + b.y}; }
F2D_HD Vec2 sub(Vec2 a, Vec2 b) { return {a.x - b.x, a.y - b.y}; }
F2D_HD Vec2 neg(Vec2 a) { return {-a.x, -a.y}; }
F2D_HD Vec2 lerp(Vec2 a, Vec2 b, float t) {
  return {(1.0f - t) * a.x + t * b.x, (1.0f - t) * a.y + t * b.y};
}
F2D_HD Vec2 mulSV(float s, Vec2 v) { return {s * v.x, s * v.y}; }
F2D_HD Vec2 mulAdd(Vec2 a, float s, Vec2 b) { return {a.x + s * b.x, a.y + s * b.y}; }
F2D_HD Vec2 mulSub(Vec2 a, float s, Vec2 b) { return {a.x - s * b.x, a.y - s * b.y}; }
F2D_HD Vec2 vmin(Vec2 a, Vec2 b) { return {minf(a.x, b.x), minf(a.y, b.y)}; }
F2D_HD Vec2 vmax(Vec2 a, Vec2 b) { return {maxf(a.x, b.x), maxf(a.y, b.y)}; }
F2D_HD float length(Vec2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
F2D_HD float lengthSq(Vec2 v) { return v.x * v.x + v.y * v.y; }
F2D_HD float distanceSq(Vec2 a, Vec2 b) {
  Vec2 c = {b.x - a.x, b.y - a.y};
  return c.x * c.x + c.y * c.y;
}

// b2Normalize, I/math_functions.h:300-311
F2D_HD Vec2 normalize(Vec2 v) {
  float len = sqrtf(v.x * v.x + v.y * v.y);
  if (len < kFltEpsilon) return {0.0f, 0.0f};
  float inv = 1.0f / len;
  return {inv * v.x, inv * v.y};
}

// b2GetLengthAndNormalize, I/math_functions.h:318-329
F2D_HD Vec2 lengthAndNormalize(float* len, Vec2 v) {
  *len = sqrtf(v.x * v.x + v.y * v.y);
  if (*len < kFltEpsilon) return {0.0f, 0.0f};
  float inv = 1.0f / *len;
  return {inv * v.x, inv * v.y};
}

// b2NormalizeRot, I/math_functions.h:332-338 (the reference compares mag > 0.0 in
// double; for a float mag that is the same predicate)
F2D_HD Rot normalizeRot(Rot q) {
  float mag = sqrtf(q.s * q.s + q.c * q.c);
  float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
  return {q.c * inv, q.s * inv};
}

// b2IntegrateRotation, I/math_functions.h:343-354
F2D_HD Rot integrateRotation(Rot q1, float da) {
  Rot q2 = {q1.c - da * q1.s, q1.s + da * q1.c};
  float mag = sqrtf(q2.s * q2.s + q2.c * q2.c);
  float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
  return {q2.c * inv, q2.s * inv};
}

// b2MulRot :445-455, b2InvMulRot :458-468
F2D_HD Rot mulRot(Rot q, Rot r) {
  Rot qr;
  qr.s = q.s * r.c + q.c * r.s;
  qr.c = q.c * r.c - q.s * r.s;
  return qr;
}
F2D_HD Rot invMulRot(Rot q, Rot r) {
  Rot qr;
  qr.s = q.c * r.s - q.s * r.c;
  qr.c = q.c * r.c + q.s * r.s;
  return qr;
}
// b2RotateVector :488-491, b2InvRotateVector :494-497
F2D_HD Vec2 rotate(Rot q, Vec2 v) { return {q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y}; }
F2D_HD Vec2 invRotate(Rot q, Vec2 v) { return {q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y}; }

// b2TransformPoint :500-506 (parenthesised sum first, then translation)
F2D_HD Vec2 transformPoint(Xf t, Vec2 p) {
  float x = (t.q.c * p.x - t.q.s * p.y) + t.p.x;
  float y = (t.q.s * p.x + t.q.c * p.y) + t.p.y;
  return {x, y};
}
// b2InvMulTransforms :532-538
F2D_HD Xf invMulTransforms(Xf A, Xf B) {
  Xf C;
  C.q = invMulRot(A.q, B.q);
  C.p = invRotate(A.q, sub(B.p, A.p));
  return C;
}

// b2AABB_Overlaps :618-622
F2D_HD bool aabbOverlaps(AABB a, AABB b) {
  return !(b.lo.x > a.hi.x || b.lo.y > a.hi.y || a.lo.x > b.hi.x || a.lo.y > b.hi.y);
}
// b2AABB_Contains :582-590
F2D_HD bool aabbContains(AABB a, AABB b) {
  bool s = true;
  s = s && a.lo.x <= b.lo.x;
  s = s && a.lo.y <= b.lo.y;
  s = s && b.hi.x <= a.hi.x;
  s = s && b.hi.y <= a.hi.y;
  return s;
}

// b2MakeSoft, S/solver.h:140-182
struct Softness {
  float biasRate, massScale, impulseScale;
};
F2D_HD Softness makeSoft(float hertz, float zeta, float h) {
  if (hertz == 0.0f) return {0.0f, 0.0f, 0.0f};
  float omega = 2.0f * kPi * hertz;
  float a1 = 2.0f * zeta + h * omega;
  float a2 = h * omega * a1;
  float a3 = 1.0f / (1.0f + a2);
  return {omega / a1, a2 * a3, a3};
}

}  // namespace f2d
