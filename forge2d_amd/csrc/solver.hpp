// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Contact and body solver stages, one constraint (or body) per lane.
//
// Coloured constraints follow the 4-wide SIMD solver's per-lane arithmetic
// (S/contact_solver.c:1438-2120): b2MulAddW(a,b,c) = a + b*c, b2MulSubW = a - b*c,
// b2MaxW/b2MinW = SSE max/min (a > b ? a : b, a < b ? a : b), blends on the sign
// of s > 0. Overflow constraints follow the scalar solver (:24-509), whose
// expression order differs (impulse formula, minimum bias velocity, clamps).
// Body integration follows S/solver.c:65-198 and finalize :543-726.
#pragma once

#include "layout.hpp"

namespace f2d {

#if defined(__HIP_DEVICE_COMPILE__)
F2D_HD float asFloat(int v) { return __int_as_float(v); }
F2D_HD int asInt(float v) { return __float_as_int(v); }
#else
inline float asFloat(int v) {
  float f;
  __builtin_memcpy(&f, &v, 4);
  return f;
}
inline int asInt(float v) {
  int i;
  __builtin_memcpy(&i, &v, 4);
  return i;
}
#endif

struct Cons {
  float* f;
  long stride;
  F2D_HD float& operator()(int field, long lane) const { return f[field * stride + lane]; }
  F2D_HD int geti(int field, long lane) const { return asInt(f[field * stride + lane]); }
  F2D_HD void seti(int field, long lane, int v) const { f[field * stride + lane] = asFloat(v); }
};

F2D_HD BodyState identityState() {
  BodyState s;
  s.v = {0.0f, 0.0f};
  s.w = 0.0f;
  s.flags = 0;
  s.dp = {0.0f, 0.0f};
  s.dq = {1.0f, 0.0f};
  return s;
}

// ---------------------------------------------------------------------------- prepare

// b2PrepareContactsTask lane body, S/contact_solver.c:1462-1627 (identical to the
// overflow prepare :50-157 for the quantities both compute).
F2D_HD void prepareContactLane(const Cons& c, long lane, const Contact& cs, int slot, int indexA, int indexB,
                               float mA, float iA, float mB, float iB, const BodyState* states,
                               Softness contactSoftness, Softness staticSoftness, float warmStartScale) {
  const Manifold& m = cs.m;
  c.seti(CN_INDEX_A, lane, indexA);
  c.seti(CN_INDEX_B, lane, indexB);
  c.seti(CN_SLOT, lane, slot);
  c.seti(CN_POINT_COUNT, lane, m.pointCount);
  Vec2 vA = {0.0f, 0.0f};
  float wA = 0.0f;
  if (indexA >= 0) {
    vA = states[indexA].v;
    wA = states[indexA].w;
  }
  Vec2 vB = {0.0f, 0.0f};
  float wB = 0.0f;
  if (indexB >= 0) {
    vB = states[indexB].v;
    wB = states[indexB].w;
  }
  c(CN_INV_MASS_A, lane) = mA;
  c(CN_INV_MASS_B, lane) = mB;
  c(CN_INV_I_A, lane) = iA;
  c(CN_INV_I_B, lane) = iB;
  {
    float k = iA + iB;
    c(CN_ROLLING_MASS, lane) = k > 0.0f ? 1.0f / k : 0.0f;
  }
  Softness soft = (indexA < 0 || indexB < 0) ? staticSoftness : contactSoftness;
  Vec2 normal = m.normal;
  c(CN_NX, lane) = normal.x;
  c(CN_NY, lane) = normal.y;
  c(CN_FRICTION, lane) = cs.friction;
  c(CN_TANGENT_SPEED, lane) = cs.tangentSpeed;
  c(CN_RESTITUTION, lane) = cs.restitution;
  c(CN_ROLLING_RES, lane) = cs.rollingResistance;
  c(CN_ROLLING_IMPULSE, lane) = warmStartScale * m.rollingImpulse;
  c(CN_BIAS_RATE, lane) = soft.biasRate;
  c(CN_MASS_SCALE, lane) = soft.massScale;
  c(CN_IMPULSE_SCALE, lane) = soft.impulseScale;
  Vec2 tangent = rightPerp(normal);
  for (int p = 0; p < 2; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    const int rvf = p == 0 ? CN_REL_VEL1 : CN_REL_VEL2;
    if (p < m.pointCount) {
      const ManifoldPoint& mp = m.points[p];
      Vec2 rA = mp.anchorA;
      Vec2 rB = mp.anchorB;
      c(CN_A1X + o, lane) = rA.x;
      c(CN_A1Y + o, lane) = rA.y;
      c(CN_B1X + o, lane) = rB.x;
      c(CN_B1Y + o, lane) = rB.y;
      c(CN_BASE_SEP1 + o, lane) = mp.separation - dot(sub(rB, rA), normal);
      c(CN_NI1 + o, lane) = warmStartScale * mp.normalImpulse;
      c(CN_TI1 + o, lane) = warmStartScale * mp.tangentImpulse;
      c(CN_TNI1 + o, lane) = 0.0f;
      float rnA = cross(rA, normal);
      float rnB = cross(rB, normal);
      float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
      c(CN_NM1 + o, lane) = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
      float rtA = cross(rA, tangent);
      float rtB = cross(rB, tangent);
      float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
      c(CN_TM1 + o, lane) = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
      Vec2 vrA = add(vA, crossSV(wA, rA));
      Vec2 vrB = add(vB, crossSV(wB, rB));
      c(rvf, lane) = dot(normal, sub(vrB, vrA));
    } else {
      c(CN_A1X + o, lane) = 0.0f;
      c(CN_A1Y + o, lane) = 0.0f;
      c(CN_B1X + o, lane) = 0.0f;
      c(CN_B1Y + o, lane) = 0.0f;
      c(CN_BASE_SEP1 + o, lane) = 0.0f;
      c(CN_NI1 + o, lane) = 0.0f;
      c(CN_TI1 + o, lane) = 0.0f;
      c(CN_TNI1 + o, lane) = 0.0f;
      c(CN_NM1 + o, lane) = 0.0f;
      c(CN_TM1 + o, lane) = 0.0f;
      c(rvf, lane) = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------- coloured (SIMD lane order)

F2D_HD BodyState gather(const BodyState* states, int idx) { return idx >= 0 ? states[idx] : identityState(); }
F2D_HD void scatterVel(BodyState* states, int idx, const BodyState& b) {
  if (idx >= 0) {
    states[idx].v = b.v;
    states[idx].w = b.w;
  }
}

// b2WarmStartContactsTask lane, S/contact_solver.c:1671-1727
F2D_HD void warmStartLane(const Cons& c, long lane, BodyState* states) {
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState bA = gather(states, ia);
  BodyState bB = gather(states, ib);
  float mA = c(CN_INV_MASS_A, lane), mB = c(CN_INV_MASS_B, lane);
  float iA = c(CN_INV_I_A, lane), iB = c(CN_INV_I_B, lane);
  float nx = c(CN_NX, lane), ny = c(CN_NY, lane);
  float tx = ny;
  float ty = 0.0f - nx;
  for (int p = 0; p < 2; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    float rAx = c(CN_A1X + o, lane), rAy = c(CN_A1Y + o, lane);
    float rBx = c(CN_B1X + o, lane), rBy = c(CN_B1Y + o, lane);
    float ni = c(CN_NI1 + o, lane), ti = c(CN_TI1 + o, lane);
    float Px = ni * nx + ti * tx;
    float Py = ni * ny + ti * ty;
    bA.w = bA.w - iA * (rAx * Py - rAy * Px);
    bA.v.x = bA.v.x - mA * Px;
    bA.v.y = bA.v.y - mA * Py;
    bB.w = bB.w + iB * (rBx * Py - rBy * Px);
    bB.v.x = bB.v.x + mB * Px;
    bB.v.y = bB.v.y + mB * Py;
  }
  float ri = c(CN_ROLLING_IMPULSE, lane);
  bA.w = bA.w - iA * ri;
  bB.w = bB.w + iB * ri;
  scatterVel(states, ia, bA);
  scatterVel(states, ib, bB);
}

// b2SolveContactsTask lane, S/contact_solver.c:1729-1967 (useBias = solve, !useBias = relax)
F2D_HD void solveLane(const Cons& c, long lane, BodyState* states, float inv_h, float minBiasVel, bool useBias) {
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState bA = gather(states, ia);
  BodyState bB = gather(states, ib);
  float mA = c(CN_INV_MASS_A, lane), mB = c(CN_INV_MASS_B, lane);
  float iA = c(CN_INV_I_A, lane), iB = c(CN_INV_I_B, lane);
  float nx = c(CN_NX, lane), ny = c(CN_NY, lane);
  float biasRate, massScale, impulseScale;
  if (useBias) {
    biasRate = c(CN_BIAS_RATE, lane);
    massScale = c(CN_MASS_SCALE, lane);
    impulseScale = c(CN_IMPULSE_SCALE, lane);
  } else {
    biasRate = 0.0f;
    massScale = 1.0f;
    impulseScale = 0.0f;
  }
  float totalNormalImpulse = 0.0f;
  float dpx = bB.dp.x - bA.dp.x;
  float dpy = bB.dp.y - bA.dp.y;
  for (int p = 0; p < 2; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    float rAx = c(CN_A1X + o, lane), rAy = c(CN_A1Y + o, lane);
    float rBx = c(CN_B1X + o, lane), rBy = c(CN_B1Y + o, lane);
    // moving anchors for current separation
    float rsAx = bA.dq.c * rAx - bA.dq.s * rAy;
    float rsAy = bA.dq.s * rAx + bA.dq.c * rAy;
    float rsBx = bB.dq.c * rBx - bB.dq.s * rBy;
    float rsBy = bB.dq.s * rBx + bB.dq.c * rBy;
    float dsx = dpx + (rsBx - rsAx);
    float dsy = dpy + (rsBy - rsAy);
    float s = (nx * dsx + ny * dsy) + c(CN_BASE_SEP1 + o, lane);
    bool mask = s > 0.0f;
    float specBias = s * inv_h;
    float softBias = maxf(biasRate * s, minBiasVel);
    float bias = mask ? specBias : softBias;
    float pointMassScale = mask ? 1.0f : massScale;
    float pointImpulseScale = mask ? 0.0f : impulseScale;
    float dvx = (bB.v.x - bB.w * rBy) - (bA.v.x - bA.w * rAy);
    float dvy = (bB.v.y + bB.w * rBx) - (bA.v.y + bA.w * rAx);
    float vn = dvx * nx + dvy * ny;
    float ni = c(CN_NI1 + o, lane);
    float negImpulse = c(CN_NM1 + o, lane) * (pointMassScale * (vn + bias)) + pointImpulseScale * ni;
    float newImpulse = maxf(ni - negImpulse, 0.0f);
    float impulse = newImpulse - ni;
    c(CN_NI1 + o, lane) = newImpulse;
    c(CN_TNI1 + o, lane) = c(CN_TNI1 + o, lane) + newImpulse;
    totalNormalImpulse = totalNormalImpulse + newImpulse;
    float Px = impulse * nx;
    float Py = impulse * ny;
    bA.v.x = bA.v.x - mA * Px;
    bA.v.y = bA.v.y - mA * Py;
    bA.w = bA.w - iA * (rAx * Py - rAy * Px);
    bB.v.x = bB.v.x + mB * Px;
    bB.v.y = bB.v.y + mB * Py;
    bB.w = bB.w + iB * (rBx * Py - rBy * Px);
  }
  float tx = ny;
  float ty = 0.0f - nx;
  float friction = c(CN_FRICTION, lane);
  float tangentSpeed = c(CN_TANGENT_SPEED, lane);
  for (int p = 0; p < 2; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    float rAx = c(CN_A1X + o, lane), rAy = c(CN_A1Y + o, lane);
    float rBx = c(CN_B1X + o, lane), rBy = c(CN_B1Y + o, lane);
    float dvx = (bB.v.x - bB.w * rBy) - (bA.v.x - bA.w * rAy);
    float dvy = (bB.v.y + bB.w * rBx) - (bA.v.y + bA.w * rAx);
    float vt = dvx * tx + dvy * ty;
    vt = vt - tangentSpeed;
    float negImpulse = c(CN_TM1 + o, lane) * vt;
    float maxFriction = friction * c(CN_NI1 + o, lane);
    float ti = c(CN_TI1 + o, lane);
    float newImpulse = ti - negImpulse;
    newImpulse = maxf(0.0f - maxFriction, minf(newImpulse, maxFriction));
    float impulse = newImpulse - ti;
    c(CN_TI1 + o, lane) = newImpulse;
    float Px = impulse * tx;
    float Py = impulse * ty;
    bA.v.x = bA.v.x - mA * Px;
    bA.v.y = bA.v.y - mA * Py;
    bA.w = bA.w - iA * (rAx * Py - rAy * Px);
    bB.v.x = bB.v.x + mB * Px;
    bB.v.y = bB.v.y + mB * Py;
    bB.w = bB.w + iB * (rBx * Py - rBy * Px);
  }
  {
    float deltaLambda = c(CN_ROLLING_MASS, lane) * (bA.w - bB.w);
    float lambda = c(CN_ROLLING_IMPULSE, lane);
    float maxLambda = c(CN_ROLLING_RES, lane) * totalNormalImpulse;
    // b2SymClampW: max(-b, min(a, b)) with -b formed by flipping the sign bit
    float a = lambda + deltaLambda;
    float nb = -maxLambda;
    float ri = maxf(nb, minf(a, maxLambda));
    c(CN_ROLLING_IMPULSE, lane) = ri;
    deltaLambda = ri - lambda;
    bA.w = bA.w - iA * deltaLambda;
    bB.w = bB.w + iB * deltaLambda;
  }
  scatterVel(states, ia, bA);
  scatterVel(states, ib, bB);
}

// b2ApplyRestitutionTask lane, S/contact_solver.c:1969-2076. A lane with zero
// restitution is skipped: in the 4-wide code such a lane is masked to zero mass,
// which leaves its impulses and velocities unchanged.
F2D_HD void restitutionLane(const Cons& c, long lane, BodyState* states, float threshold) {
  float restitution = c(CN_RESTITUTION, lane);
  if (restitution == 0.0f) return;
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState bA = gather(states, ia);
  BodyState bB = gather(states, ib);
  float mA = c(CN_INV_MASS_A, lane), mB = c(CN_INV_MASS_B, lane);
  float iA = c(CN_INV_I_A, lane), iB = c(CN_INV_I_B, lane);
  float nx = c(CN_NX, lane), ny = c(CN_NY, lane);
  for (int p = 0; p < 2; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    const int rvf = p == 0 ? CN_REL_VEL1 : CN_REL_VEL2;
    float relVel = c(rvf, lane);
    bool mask1 = (relVel + threshold) > 0.0f;
    bool mask2 = c(CN_TNI1 + o, lane) == 0.0f;
    bool mask = mask1 || mask2;
    float mass = mask ? 0.0f : c(CN_NM1 + o, lane);
    float rAx = c(CN_A1X + o, lane), rAy = c(CN_A1Y + o, lane);
    float rBx = c(CN_B1X + o, lane), rBy = c(CN_B1Y + o, lane);
    float dvx = (bB.v.x - bB.w * rBy) - (bA.v.x - bA.w * rAy);
    float dvy = (bB.v.y + bB.w * rBx) - (bA.v.y + bA.w * rAx);
    float vn = dvx * nx + dvy * ny;
    float negImpulse = mass * (vn + restitution * relVel);
    float ni = c(CN_NI1 + o, lane);
    float newImpulse = maxf(ni - negImpulse, 0.0f);
    float impulse = newImpulse - ni;
    c(CN_NI1 + o, lane) = newImpulse;
    float Px = impulse * nx;
    float Py = impulse * ny;
    bA.v.x = bA.v.x - mA * Px;
    bA.v.y = bA.v.y - mA * Py;
    bA.w = bA.w - iA * (rAx * Py - rAy * Px);
    bB.v.x = bB.v.x + mB * Px;
    bB.v.y = bB.v.y + mB * Py;
    bB.w = bB.w + iB * (rBx * Py - rBy * Px);
  }
  scatterVel(states, ia, bA);
  scatterVel(states, ib, bB);
}

// b2StoreImpulsesTask lane, S/contact_solver.c:2078-2120 (writes both points)
F2D_HD void storeLane(const Cons& c, long lane, Contact& cs) {
  Manifold& m = cs.m;
  m.rollingImpulse = c(CN_ROLLING_IMPULSE, lane);
  m.points[0].normalImpulse = c(CN_NI1, lane);
  m.points[0].tangentImpulse = c(CN_TI1, lane);
  m.points[0].totalNormalImpulse = c(CN_TNI1, lane);
  m.points[0].normalVelocity = c(CN_REL_VEL1, lane);
  m.points[1].normalImpulse = c(CN_NI2, lane);
  m.points[1].tangentImpulse = c(CN_TI2, lane);
  m.points[1].totalNormalImpulse = c(CN_TNI2, lane);
  m.points[1].normalVelocity = c(CN_REL_VEL2, lane);
}

// ---------------------------------------------------------------------------- overflow (scalar order)

// Static bodies have no solver state: the scalar solver reads and writes a dummy
// identity state for them (S/contact_solver.c:175-176). Their inverse mass and
// inertia are zero, so for finite impulses the dummy's velocities stay exactly +0
// and its deltas the identity; reading the identity and dropping the write is
// therefore the same computation.
F2D_HD BodyState loadState(const BodyState* states, int idx) { return idx < 0 ? identityState() : states[idx]; }
F2D_HD void storeVel(BodyState* states, int idx, Vec2 v, float w) {
  if (idx >= 0) {
    states[idx].v = v;
    states[idx].w = w;
  }
}

// b2WarmStartOverflowContacts body, S/contact_solver.c:160-225
F2D_HD void overflowWarmStart(const Cons& c, long lane, BodyState* states) {
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  float mA = c(CN_INV_MASS_A, lane), iA = c(CN_INV_I_A, lane);
  float mB = c(CN_INV_MASS_B, lane), iB = c(CN_INV_I_B, lane);
  Vec2 normal = {c(CN_NX, lane), c(CN_NY, lane)};
  Vec2 tangent = rightPerp(normal);
  int pointCount = c.geti(CN_POINT_COUNT, lane);
  for (int p = 0; p < pointCount; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    Vec2 rA = {c(CN_A1X + o, lane), c(CN_A1Y + o, lane)};
    Vec2 rB = {c(CN_B1X + o, lane), c(CN_B1Y + o, lane)};
    Vec2 P = add(mulSV(c(CN_NI1 + o, lane), normal), mulSV(c(CN_TI1 + o, lane), tangent));
    wA -= iA * cross(rA, P);
    vA = mulAdd(vA, -mA, P);
    wB += iB * cross(rB, P);
    vB = mulAdd(vB, mB, P);
  }
  float ri = c(CN_ROLLING_IMPULSE, lane);
  wA -= iA * ri;
  wB += iB * ri;
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// b2SolveOverflowContacts body, S/contact_solver.c:227-380
F2D_HD void overflowSolve(const Cons& c, long lane, BodyState* states, float inv_h, float pushout, bool useBias) {
  float mA = c(CN_INV_MASS_A, lane), iA = c(CN_INV_I_A, lane);
  float mB = c(CN_INV_MASS_B, lane), iB = c(CN_INV_I_B, lane);
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState sA = loadState(states, ia);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Rot dqA = sA.dq;
  BodyState sB = loadState(states, ib);
  Vec2 vB = sB.v;
  float wB = sB.w;
  Rot dqB = sB.dq;
  Vec2 dp = sub(sB.dp, sA.dp);
  Vec2 normal = {c(CN_NX, lane), c(CN_NY, lane)};
  Vec2 tangent = rightPerp(normal);
  float friction = c(CN_FRICTION, lane);
  Softness softness = {c(CN_BIAS_RATE, lane), c(CN_MASS_SCALE, lane), c(CN_IMPULSE_SCALE, lane)};
  int pointCount = c.geti(CN_POINT_COUNT, lane);
  float totalNormalImpulse = 0.0f;
  for (int p = 0; p < pointCount; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    Vec2 rA = {c(CN_A1X + o, lane), c(CN_A1Y + o, lane)};
    Vec2 rB = {c(CN_B1X + o, lane), c(CN_B1Y + o, lane)};
    Vec2 ds = add(dp, sub(rotate(dqB, rB), rotate(dqA, rA)));
    float s = c(CN_BASE_SEP1 + o, lane) + dot(ds, normal);
    float velocityBias = 0.0f;
    float massScale = 1.0f;
    float impulseScale = 0.0f;
    if (s > 0.0f) {
      velocityBias = s * inv_h;
    } else if (useBias) {
      velocityBias = maxf(softness.biasRate * s, -pushout);
      massScale = softness.massScale;
      impulseScale = softness.impulseScale;
    }
    Vec2 vrA = add(vA, crossSV(wA, rA));
    Vec2 vrB = add(vB, crossSV(wB, rB));
    float vn = dot(sub(vrB, vrA), normal);
    float ni = c(CN_NI1 + o, lane);
    float impulse = -c(CN_NM1 + o, lane) * massScale * (vn + velocityBias) - impulseScale * ni;
    float newImpulse = maxf(ni + impulse, 0.0f);
    impulse = newImpulse - ni;
    c(CN_NI1 + o, lane) = newImpulse;
    c(CN_TNI1 + o, lane) = c(CN_TNI1 + o, lane) + newImpulse;
    totalNormalImpulse += newImpulse;
    Vec2 P = mulSV(impulse, normal);
    vA = mulSub(vA, mA, P);
    wA -= iA * cross(rA, P);
    vB = mulAdd(vB, mB, P);
    wB += iB * cross(rB, P);
  }
  float tangentSpeed = c(CN_TANGENT_SPEED, lane);
  for (int p = 0; p < pointCount; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    Vec2 rA = {c(CN_A1X + o, lane), c(CN_A1Y + o, lane)};
    Vec2 rB = {c(CN_B1X + o, lane), c(CN_B1Y + o, lane)};
    Vec2 vrB = add(vB, crossSV(wB, rB));
    Vec2 vrA = add(vA, crossSV(wA, rA));
    float vt = dot(sub(vrB, vrA), tangent) - tangentSpeed;
    float impulse = c(CN_TM1 + o, lane) * (-vt);
    float maxFriction = friction * c(CN_NI1 + o, lane);
    float ti = c(CN_TI1 + o, lane);
    float newImpulse = clampf(ti + impulse, -maxFriction, maxFriction);
    impulse = newImpulse - ti;
    c(CN_TI1 + o, lane) = newImpulse;
    Vec2 P = mulSV(impulse, tangent);
    vA = mulSub(vA, mA, P);
    wA -= iA * cross(rA, P);
    vB = mulAdd(vB, mB, P);
    wB += iB * cross(rB, P);
  }
  {
    float deltaLambda = -c(CN_ROLLING_MASS, lane) * (wB - wA);
    float lambda = c(CN_ROLLING_IMPULSE, lane);
    float maxLambda = c(CN_ROLLING_RES, lane) * totalNormalImpulse;
    float ri = clampf(lambda + deltaLambda, -maxLambda, maxLambda);
    c(CN_ROLLING_IMPULSE, lane) = ri;
    deltaLambda = ri - lambda;
    wA -= iA * deltaLambda;
    wB += iB * deltaLambda;
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// b2ApplyOverflowRestitution body, S/contact_solver.c:382-478
F2D_HD void overflowRestitution(const Cons& c, long lane, BodyState* states, float threshold) {
  float restitution = c(CN_RESTITUTION, lane);
  if (restitution == 0.0f) return;
  float mA = c(CN_INV_MASS_A, lane), iA = c(CN_INV_I_A, lane);
  float mB = c(CN_INV_MASS_B, lane), iB = c(CN_INV_I_B, lane);
  int ia = c.geti(CN_INDEX_A, lane), ib = c.geti(CN_INDEX_B, lane);
  BodyState sA = loadState(states, ia);
  Vec2 vA = sA.v;
  float wA = sA.w;
  BodyState sB = loadState(states, ib);
  Vec2 vB = sB.v;
  float wB = sB.w;
  Vec2 normal = {c(CN_NX, lane), c(CN_NY, lane)};
  int pointCount = c.geti(CN_POINT_COUNT, lane);
  for (int p = 0; p < pointCount; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    const int rvf = p == 0 ? CN_REL_VEL1 : CN_REL_VEL2;
    float relVel = c(rvf, lane);
    if (relVel > -threshold || c(CN_TNI1 + o, lane) == 0.0f) continue;
    Vec2 rA = {c(CN_A1X + o, lane), c(CN_A1Y + o, lane)};
    Vec2 rB = {c(CN_B1X + o, lane), c(CN_B1Y + o, lane)};
    Vec2 vrB = add(vB, crossSV(wB, rB));
    Vec2 vrA = add(vA, crossSV(wA, rA));
    float vn = dot(sub(vrB, vrA), normal);
    float impulse = -c(CN_NM1 + o, lane) * (vn + restitution * relVel);
    float ni = c(CN_NI1 + o, lane);
    float newImpulse = maxf(ni + impulse, 0.0f);
    impulse = newImpulse - ni;
    c(CN_NI1 + o, lane) = newImpulse;
    c(CN_TNI1 + o, lane) = c(CN_TNI1 + o, lane) + impulse;
    Vec2 P = mulSV(impulse, normal);
    vA = mulSub(vA, mA, P);
    wA -= iA * cross(rA, P);
    vB = mulAdd(vB, mB, P);
    wB += iB * cross(rB, P);
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// b2StoreOverflowImpulses body, S/contact_solver.c:480-509 (writes pointCount points)
F2D_HD void overflowStore(const Cons& c, long lane, Contact& cs) {
  Manifold& m = cs.m;
  for (int p = 0; p < m.pointCount; ++p) {
    const int o = p == 0 ? 0 : (CN_A2X - CN_A1X);
    const int rvf = p == 0 ? CN_REL_VEL1 : CN_REL_VEL2;
    m.points[p].normalImpulse = c(CN_NI1 + o, lane);
    m.points[p].tangentImpulse = c(CN_TI1 + o, lane);
    m.points[p].totalNormalImpulse = c(CN_TNI1 + o, lane);
    m.points[p].normalVelocity = c(rvf, lane);
  }
  m.rollingImpulse = c(CN_ROLLING_IMPULSE, lane);
}

// ---------------------------------------------------------------------------- bodies

// b2IntegrateVelocitiesTask body, S/solver.c:65-129
F2D_HD void integrateVelocity(BodySim& sim, BodyState& state, Vec2 gravity, float h, float maxLinearSpeed,
                              float maxAngularSpeed) {
  float maxLinearSpeedSquared = maxLinearSpeed * maxLinearSpeed;
  float maxAngularSpeedSquared = maxAngularSpeed * maxAngularSpeed;
  Vec2 v = state.v;
  float w = state.w;
  float linearDamping = 1.0f / (1.0f + h * sim.linearDamping);
  float angularDamping = 1.0f / (1.0f + h * sim.angularDamping);
  float gravityScale = sim.invMass > 0.0f ? sim.gravityScale : 0.0f;
  Vec2 linearVelocityDelta = add(mulSV(h * sim.invMass, sim.force), mulSV(h * gravityScale, gravity));
  float angularVelocityDelta = h * sim.invI * sim.torque;
  v = mulAdd(linearVelocityDelta, linearDamping, v);
  w = angularVelocityDelta + angularDamping * w;
  if (dot(v, v) > maxLinearSpeedSquared) {
    float ratio = maxLinearSpeed / length(v);
    v = mulSV(ratio, v);
    sim.flags |= BF_CAPPED_STEP;
  }
  if (w * w > maxAngularSpeedSquared && (sim.flags & BF_ALLOW_FAST_ROTATION) == 0) {
    float ratio = maxAngularSpeed / absf(w);
    w *= ratio;
    sim.flags |= BF_CAPPED_STEP;
  }
  state.v = v;
  state.w = w;
}

// The per-step body terms of b2IntegrateVelocitiesTask (S/solver.c:65-129): the
// terms that do not change across substeps, evaluated once per step with the
// reference's own expressions, so each substep's update is bit-identical to
// recomputing them.
enum : int { BC_STATIC = 1, BC_FAST_ROTATION = 2, BC_CAPPED = 4 };
struct alignas(16) BodyConst {
  Vec2 dv;        // h*invMass*force + h*gravityScale*gravity
  float dw;       // h*invI*torque
  float linDamp;  // 1 / (1 + h*linearDamping)
  float angDamp;  // 1 / (1 + h*angularDamping)
  float invMass, invI;
  int flags;      // BC_*
};

F2D_HD BodyConst makeBodyConst(const BodySim& sim, Vec2 gravity, float h) {
  BodyConst k;
  k.linDamp = 1.0f / (1.0f + h * sim.linearDamping);
  k.angDamp = 1.0f / (1.0f + h * sim.angularDamping);
  float gravityScale = sim.invMass > 0.0f ? sim.gravityScale : 0.0f;
  k.dv = add(mulSV(h * sim.invMass, sim.force), mulSV(h * gravityScale, gravity));
  k.dw = h * sim.invI * sim.torque;
  k.invMass = sim.invMass;
  k.invI = sim.invI;
  // bodies outside the awake set (static or sleeping) take no solver updates
  k.flags = (sim.setIndex != kAwakeSet ? BC_STATIC : 0) | ((sim.flags & BF_ALLOW_FAST_ROTATION) ? BC_FAST_ROTATION : 0);
  return k;
}

// b2IntegrateVelocitiesTask body from the per-step terms
F2D_HD void integrateVelocityConst(BodyConst& k, BodyState& state, float maxLinearSpeed, float maxAngularSpeed) {
  float maxLinearSpeedSquared = maxLinearSpeed * maxLinearSpeed;
  float maxAngularSpeedSquared = maxAngularSpeed * maxAngularSpeed;
  Vec2 v = mulAdd(k.dv, k.linDamp, state.v);
  float w = k.dw + k.angDamp * state.w;
  if (dot(v, v) > maxLinearSpeedSquared) {
    float ratio = maxLinearSpeed / length(v);
    v = mulSV(ratio, v);
    k.flags |= BC_CAPPED;
  }
  if (w * w > maxAngularSpeedSquared && (k.flags & BC_FAST_ROTATION) == 0) {
    float ratio = maxAngularSpeed / absf(w);
    w *= ratio;
    k.flags |= BC_CAPPED;
  }
  state.v = v;
  state.w = w;
}

// b2IntegratePositionsTask body, S/solver.c:182-198
F2D_HD void integratePosition(BodyState& state, float h) {
  state.dq = integrateRotation(state.dq, h * state.w);
  state.dp = mulAdd(state.dp, h, state.v);
}

}  // namespace f2d
