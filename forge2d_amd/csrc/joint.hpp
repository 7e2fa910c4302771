// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Revolute, prismatic, weld, wheel and distance joint stages, one joint per lane, in the
// reference's operation order: revolute prepare S/joint.c:1348-1355 +
// S/revolute_joint.c:202-259, warm start :261-287, solve :289-465; the others below.
// Static bodies read the identity body state as in the reference.
#pragma once

#include "solver.hpp"

namespace f2d {

// b2Atan2, S/math_functions.c:61-101 (deterministic minimax approximation)
F2D_HD float atan2Approx(float y, float x) {
  if (x == 0.0f && y == 0.0f) return 0.0f;
  float ax = absf(x);
  float ay = absf(y);
  float mx = maxf(ay, ax);
  float mn = minf(ay, ax);
  float a = mn / mx;
  float s = a * a;
  float c = s * a;
  float q = s * s;
  float r = 0.024840285f * q + 0.18681418f;
  float t = -0.094097948f * q - 0.33213072f;
  r = r * s + t;
  r = r * c + a;
  if (ay > ax) r = 1.57079637f - r;
  if (x < 0) r = 3.14159274f - r;
  if (y < 0) r = -r;
  return r;
}

// b2RelativeAngle, I/math_functions.h:472-479
F2D_HD float relativeAngle(Rot b, Rot a) {
  float s = b.s * a.c - b.c * a.s;
  float c = b.c * a.c + b.s * a.s;
  return atan2Approx(s, c);
}

// b2UnwindAngle, I/math_functions.h:481-485
F2D_HD float unwindAngle(float radians) { return remainderf(radians, 2.0f * kPi); }

// b2Solve22, I/math_functions.h:569-579
F2D_HD Vec2 solve22(float a11, float a12, float a21, float a22, Vec2 b) {
  float det = a11 * a22 - a12 * a21;
  if (det != 0.0f) det = 1.0f / det;
  return {det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x)};
}

F2D_HD void prepareRevolute(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  // constraint softness, S/joint.c:1350-1352
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  j.invMassA = a.invMass;
  j.invMassB = b.invMass;
  j.invIA = a.invI;
  j.invIB = b.invI;
  j.anchorA = rotate(a.xf.q, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(b.xf.q, sub(j.localOriginAnchorB, b.localCenter));
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;  // joint->indexA, S/joint.c:1380-1388
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  j.deltaCenter = sub(b.center, a.center);
  j.deltaAngle = relativeAngle(b.xf.q, a.xf.q);
  float k = j.invIA + j.invIB;
  j.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
  j.springSoftness = makeSoft(j.hertz, j.dampingRatio, h);
  if (warmStartScale == 0.0f) {
    j.linearImpulse = {0.0f, 0.0f};
    j.springImpulse = 0.0f;
    j.motorImpulse = 0.0f;
    j.lowerImpulse = 0.0f;
    j.upperImpulse = 0.0f;
  }
}

F2D_HD void warmStartRevolute(Joint& j, BodyState* states) {
  float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  Vec2 rA = rotate(sA.dq, j.anchorA);
  Vec2 rB = rotate(sB.dq, j.anchorB);
  float axialImpulse = j.springImpulse + j.motorImpulse + j.lowerImpulse - j.upperImpulse;
  sA.v = mulSub(sA.v, mA, j.linearImpulse);
  sA.w -= iA * (cross(rA, j.linearImpulse) + axialImpulse);
  sB.v = mulAdd(sB.v, mB, j.linearImpulse);
  sB.w += iB * (cross(rB, j.linearImpulse) + axialImpulse);
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD void solveRevolute(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  const Rot dqA = sA.dq;
  const Rot dqB = sB.dq;
  bool fixedRotation = (iA + iB == 0.0f);
  if (j.enableSpring && fixedRotation == false) {
    float jointAngle = relativeAngle(sB.dq, sA.dq) + j.deltaAngle;
    float jointAngleDelta = unwindAngle(jointAngle - j.targetAngle);
    float C = jointAngleDelta;
    float bias = j.springSoftness.biasRate * C;
    float massScale = j.springSoftness.massScale;
    float impulseScale = j.springSoftness.impulseScale;
    float Cdot = wB - wA;
    float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.springImpulse;
    j.springImpulse += impulse;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  if (j.enableMotor && fixedRotation == false) {
    float Cdot = wB - wA - j.motorSpeed;
    float impulse = -j.axialMass * Cdot;
    float oldImpulse = j.motorImpulse;
    float maxImpulse = h * j.maxMotorTorque;
    j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
    impulse = j.motorImpulse - oldImpulse;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  if (j.enableLimit && fixedRotation == false) {
    float jointAngle = relativeAngle(dqB, dqA) + j.deltaAngle - j.referenceAngle;
    jointAngle = unwindAngle(jointAngle);
    {
      float C = jointAngle - j.lowerAngle;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      float Cdot = wB - wA;
      float oldImpulse = j.lowerImpulse;
      float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * oldImpulse;
      j.lowerImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.lowerImpulse - oldImpulse;
      wA -= iA * impulse;
      wB += iB * impulse;
    }
    {
      float C = j.upperAngle - jointAngle;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      float Cdot = wA - wB;
      float oldImpulse = j.upperImpulse;
      float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * oldImpulse;
      j.upperImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.upperImpulse - oldImpulse;
      wA += iA * impulse;
      wB -= iB * impulse;
    }
  }
  {
    Vec2 rA = rotate(sA.dq, j.anchorA);
    Vec2 rB = rotate(sB.dq, j.anchorB);
    Vec2 Cdot = sub(add(vB, crossSV(wB, rB)), add(vA, crossSV(wA, rA)));
    Vec2 bias = {0.0f, 0.0f};
    float massScale = 1.0f;
    float impulseScale = 0.0f;
    if (useBias) {
      Vec2 dcA = sA.dp;
      Vec2 dcB = sB.dp;
      Vec2 separation = add(add(sub(dcB, dcA), sub(rB, rA)), j.deltaCenter);
      bias = mulSV(j.constraintSoftness.biasRate, separation);
      massScale = j.constraintSoftness.massScale;
      impulseScale = j.constraintSoftness.impulseScale;
    }
    float k11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
    float k12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    float k21 = k12;
    float k22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
    // b2Mat22 {cx = (k11, k21), cy = (k12, k22)}: a11 = cx.x, a12 = cy.x, a21 = cx.y, a22 = cy.y
    Vec2 b = solve22(k11, k12, k21, k22, add(Cdot, bias));
    Vec2 impulse;
    impulse.x = -massScale * b.x - impulseScale * j.linearImpulse.x;
    impulse.y = -massScale * b.y - impulseScale * j.linearImpulse.y;
    j.linearImpulse.x += impulse.x;
    j.linearImpulse.y += impulse.y;
    vA = mulSub(vA, mA, impulse);
    wA -= iA * cross(rA, impulse);
    vB = mulAdd(vB, mB, impulse);
    wB += iB * cross(rB, impulse);
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// Prismatic joint stages, S/prismatic_joint.c: prepare :279-347 (with the common
// part of S/joint.c:1348-1355), warm start :349-392, solve :394-631.
F2D_HD_COLD void preparePrismatic(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  const float mA = a.invMass, iA = a.invI, mB = b.invMass, iB = b.invI;
  j.invMassA = mA;
  j.invMassB = mB;
  j.invIA = iA;
  j.invIB = iB;
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  const Rot qA = a.xf.q, qB = b.xf.q;
  j.anchorA = rotate(qA, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(qB, sub(j.localOriginAnchorB, b.localCenter));
  j.axisA = rotate(qA, j.localAxisA);
  j.deltaCenter = sub(b.center, a.center);
  j.deltaAngle = relativeAngle(qB, qA) - j.referenceAngle;
  j.deltaAngle = unwindAngle(j.deltaAngle);
  const Vec2 rA = j.anchorA, rB = j.anchorB;
  const Vec2 d = add(j.deltaCenter, sub(rB, rA));
  const float a1 = cross(add(d, rA), j.axisA);
  const float a2 = cross(rB, j.axisA);
  const float k = mA + mB + iA * a1 * a1 + iB * a2 * a2;
  j.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
  j.springSoftness = makeSoft(j.hertz, j.dampingRatio, h);
  if (warmStartScale == 0.0f) {
    j.linearImpulse = {0.0f, 0.0f};
    j.springImpulse = 0.0f;
    j.motorImpulse = 0.0f;
    j.lowerImpulse = 0.0f;
    j.upperImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartPrismatic(Joint& j, BodyState* states) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 d = add(add(sub(sB.dp, sA.dp), j.deltaCenter), sub(rB, rA));
  const Vec2 axisA = rotate(sA.dq, j.axisA);
  // the impulse acts at body B's anchor
  const float a1 = cross(add(d, rA), axisA);
  const float a2 = cross(rB, axisA);
  const float axialImpulse = j.springImpulse + j.motorImpulse + j.lowerImpulse - j.upperImpulse;
  const Vec2 perpA = leftPerp(axisA);
  const float s1 = cross(add(d, rA), perpA);
  const float s2 = cross(rB, perpA);
  const float perpImpulse = j.linearImpulse.x;
  const float angleImpulse = j.linearImpulse.y;
  const Vec2 P = add(mulSV(axialImpulse, axisA), mulSV(perpImpulse, perpA));
  const float LA = axialImpulse * a1 + perpImpulse * s1 + angleImpulse;
  const float LB = axialImpulse * a2 + perpImpulse * s2 + angleImpulse;
  sA.v = mulSub(sA.v, mA, P);
  sA.w -= iA * LA;
  sB.v = mulAdd(sB.v, mB, P);
  sB.w += iB * LB;
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD_COLD void solvePrismatic(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 d = add(add(sub(sB.dp, sA.dp), j.deltaCenter), sub(rB, rA));
  const Vec2 axisA = rotate(sA.dq, j.axisA);
  const float translation = dot(axisA, d);
  // torques of axial forces
  const float a1 = cross(add(d, rA), axisA);
  const float a2 = cross(rB, axisA);
  if (j.enableSpring) {
    // a real spring: applied in the relax iterations too
    const float C = translation - j.targetAngle;
    const float bias = j.springSoftness.biasRate * C;
    const float massScale = j.springSoftness.massScale;
    const float impulseScale = j.springSoftness.impulseScale;
    const float Cdot = dot(axisA, sub(vB, vA)) + a2 * wB - a1 * wA;
    const float deltaImpulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.springImpulse;
    j.springImpulse += deltaImpulse;
    const Vec2 P = mulSV(deltaImpulse, axisA);
    const float LA = deltaImpulse * a1;
    const float LB = deltaImpulse * a2;
    vA = mulSub(vA, mA, P);
    wA -= iA * LA;
    vB = mulAdd(vB, mB, P);
    wB += iB * LB;
  }
  if (j.enableMotor) {
    const float Cdot = dot(axisA, sub(vB, vA)) + a2 * wB - a1 * wA;
    float impulse = j.axialMass * (j.motorSpeed - Cdot);
    const float oldImpulse = j.motorImpulse;
    const float maxImpulse = h * j.maxMotorTorque;
    j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
    impulse = j.motorImpulse - oldImpulse;
    const Vec2 P = mulSV(impulse, axisA);
    const float LA = impulse * a1;
    const float LB = impulse * a2;
    vA = mulSub(vA, mA, P);
    wA -= iA * LA;
    vB = mulAdd(vB, mB, P);
    wB += iB * LB;
  }
  if (j.enableLimit) {
    {
      // lower limit
      const float C = translation - j.lowerAngle;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;  // speculative
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      const float oldImpulse = j.lowerImpulse;
      const float Cdot = dot(axisA, sub(vB, vA)) + a2 * wB - a1 * wA;
      float impulse = -j.axialMass * massScale * (Cdot + bias) - impulseScale * oldImpulse;
      j.lowerImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.lowerImpulse - oldImpulse;
      const Vec2 P = mulSV(impulse, axisA);
      const float LA = impulse * a1;
      const float LB = impulse * a2;
      vA = mulSub(vA, mA, P);
      wA -= iA * LA;
      vB = mulAdd(vB, mB, P);
      wB += iB * LB;
    }
    {
      // upper limit, signs flipped so that C and the impulse stay positive
      const float C = j.upperAngle - translation;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      const float oldImpulse = j.upperImpulse;
      const float Cdot = dot(axisA, sub(vA, vB)) + a1 * wA - a2 * wB;
      float impulse = -j.axialMass * massScale * (Cdot + bias) - impulseScale * oldImpulse;
      j.upperImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.upperImpulse - oldImpulse;
      const Vec2 P = mulSV(impulse, axisA);
      const float LA = impulse * a1;
      const float LB = impulse * a2;
      vA = mulAdd(vA, mA, P);
      wA += iA * LA;
      vB = mulSub(vB, mB, P);
      wB -= iB * LB;
    }
  }
  {
    // the perpendicular and angular constraints as one 2x2 block
    const Vec2 perpA = leftPerp(axisA);
    const float s1 = cross(add(d, rA), perpA);
    const float s2 = cross(rB, perpA);
    Vec2 Cdot;
    Cdot.x = dot(perpA, sub(vB, vA)) + s2 * wB - s1 * wA;
    Cdot.y = wB - wA;
    Vec2 bias = {0.0f, 0.0f};
    float massScale = 1.0f, impulseScale = 0.0f;
    if (useBias) {
      Vec2 C;
      C.x = dot(perpA, d);
      C.y = relativeAngle(sB.dq, sA.dq) + j.deltaAngle;
      bias = mulSV(j.constraintSoftness.biasRate, C);
      massScale = j.constraintSoftness.massScale;
      impulseScale = j.constraintSoftness.impulseScale;
    }
    const float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
    const float k12 = iA * s1 + iB * s2;
    float k22 = iA + iB;
    if (k22 == 0.0f) k22 = 1.0f;  // bodies with fixed rotation
    // b2Mat22 {cx = (k11, k12), cy = (k12, k22)}
    const Vec2 b = solve22(k11, k12, k12, k22, add(Cdot, bias));
    Vec2 impulse;
    impulse.x = -massScale * b.x - impulseScale * j.linearImpulse.x;
    impulse.y = -massScale * b.y - impulseScale * j.linearImpulse.y;
    j.linearImpulse.x += impulse.x;
    j.linearImpulse.y += impulse.y;
    const Vec2 P = mulSV(impulse.x, perpA);
    const float LA = impulse.x * s1 + impulse.y;
    const float LB = impulse.x * s2 + impulse.y;
    vA = mulSub(vA, mA, P);
    wA -= iA * LA;
    vB = mulAdd(vB, mB, P);
    wB += iB * LB;
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// Weld joint stages, S/weld_joint.c: prepare :91-163, warm start :165-188, solve :190-275.
F2D_HD_COLD void prepareWeld(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  const float iA = a.invI, iB = b.invI;
  j.invMassA = a.invMass;
  j.invMassB = b.invMass;
  j.invIA = iA;
  j.invIB = iB;
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  const Rot qA = a.xf.q, qB = b.xf.q;
  j.anchorA = rotate(qA, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(qB, sub(j.localOriginAnchorB, b.localCenter));
  j.deltaCenter = sub(b.center, a.center);
  j.deltaAngle = relativeAngle(qB, qA) - j.referenceAngle;
  j.deltaAngle = unwindAngle(j.deltaAngle);
  const float ka = iA + iB;
  j.axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
  j.springSoftness = j.hertz == 0.0f ? j.constraintSoftness : makeSoft(j.hertz, j.dampingRatio, h);
  j.angularSoftness =
      j.angularHertz == 0.0f ? j.constraintSoftness : makeSoft(j.angularHertz, j.angularDampingRatio, h);
  if (warmStartScale == 0.0f) {
    j.linearImpulse = {0.0f, 0.0f};
    j.springImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartWeld(Joint& j, BodyState* states) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  sA.v = mulSub(sA.v, mA, j.linearImpulse);
  sA.w -= iA * (cross(rA, j.linearImpulse) + j.springImpulse);
  sB.v = mulAdd(sB.v, mB, j.linearImpulse);
  sB.w += iB * (cross(rB, j.linearImpulse) + j.springImpulse);
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD_COLD void solveWeld(Joint& j, BodyState* states, bool useBias) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  {
    // angular constraint
    float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
    if (useBias || j.angularHertz > 0.0f) {
      const float C = relativeAngle(sB.dq, sA.dq) + j.deltaAngle;
      bias = j.angularSoftness.biasRate * C;
      massScale = j.angularSoftness.massScale;
      impulseScale = j.angularSoftness.impulseScale;
    }
    const float Cdot = wB - wA;
    const float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.springImpulse;
    j.springImpulse += impulse;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  {
    // linear constraint
    const Vec2 rA = rotate(sA.dq, j.anchorA);
    const Vec2 rB = rotate(sB.dq, j.anchorB);
    Vec2 bias = {0.0f, 0.0f};
    float massScale = 1.0f, impulseScale = 0.0f;
    if (useBias || j.hertz > 0.0f) {
      const Vec2 C = add(add(sub(sB.dp, sA.dp), sub(rB, rA)), j.deltaCenter);
      bias = mulSV(j.springSoftness.biasRate, C);
      massScale = j.springSoftness.massScale;
      impulseScale = j.springSoftness.impulseScale;
    }
    const Vec2 Cdot = sub(add(vB, crossSV(wB, rB)), add(vA, crossSV(wA, rA)));
    const float k11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
    const float k12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    const float k22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
    // b2Mat22 {cx = (k11, k12), cy = (k12, k22)}
    const Vec2 b = solve22(k11, k12, k12, k22, add(Cdot, bias));
    const Vec2 impulse = {-massScale * b.x - impulseScale * j.linearImpulse.x,
                          -massScale * b.y - impulseScale * j.linearImpulse.y};
    j.linearImpulse = add(j.linearImpulse, impulse);
    vA = mulSub(vA, mA, impulse);
    wA -= iA * cross(rA, impulse);
    vB = mulAdd(vB, mB, impulse);
    wB += iB * cross(rB, impulse);
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// Wheel joint stages, S/wheel_joint.c: prepare :183-263, warm start :265-304, solve :306-491.
F2D_HD_COLD void prepareWheel(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  const float mA = a.invMass, iA = a.invI, mB = b.invMass, iB = b.invI;
  j.invMassA = mA;
  j.invMassB = mB;
  j.invIA = iA;
  j.invIB = iB;
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  const Rot qA = a.xf.q, qB = b.xf.q;
  j.anchorA = rotate(qA, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(qB, sub(j.localOriginAnchorB, b.localCenter));
  j.axisA = rotate(qA, j.localAxisA);
  j.deltaCenter = sub(b.center, a.center);
  const Vec2 rA = j.anchorA, rB = j.anchorB;
  const Vec2 d = add(j.deltaCenter, sub(rB, rA));
  const Vec2 axisA = j.axisA;
  const Vec2 perpA = leftPerp(axisA);
  // perpendicular constraint (keeps the wheel on its line)
  const float s1 = cross(add(d, rA), perpA);
  const float s2 = cross(rB, perpA);
  const float kp = mA + mB + iA * s1 * s1 + iB * s2 * s2;
  j.perpMass = kp > 0.0f ? 1.0f / kp : 0.0f;
  // spring constraint
  const float a1 = cross(add(d, rA), axisA);
  const float a2 = cross(rB, axisA);
  const float ka = mA + mB + iA * a1 * a1 + iB * a2 * a2;
  j.axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
  j.springSoftness = makeSoft(j.hertz, j.dampingRatio, h);
  const float km = iA + iB;
  j.motorMass = km > 0.0f ? 1.0f / km : 0.0f;
  if (warmStartScale == 0.0f) {
    j.linearImpulse.x = 0.0f;
    j.springImpulse = 0.0f;
    j.motorImpulse = 0.0f;
    j.lowerImpulse = 0.0f;
    j.upperImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartWheel(Joint& j, BodyState* states) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 d = add(add(sub(sB.dp, sA.dp), j.deltaCenter), sub(rB, rA));
  const Vec2 axisA = rotate(sA.dq, j.axisA);
  const Vec2 perpA = leftPerp(axisA);
  const float a1 = cross(add(d, rA), axisA);
  const float a2 = cross(rB, axisA);
  const float s1 = cross(add(d, rA), perpA);
  const float s2 = cross(rB, perpA);
  const float axialImpulse = j.springImpulse + j.lowerImpulse - j.upperImpulse;
  const float perpImpulse = j.linearImpulse.x;
  const Vec2 P = add(mulSV(axialImpulse, axisA), mulSV(perpImpulse, perpA));
  const float LA = axialImpulse * a1 + perpImpulse * s1 + j.motorImpulse;
  const float LB = axialImpulse * a2 + perpImpulse * s2 + j.motorImpulse;
  sA.v = mulSub(sA.v, mA, P);
  sA.w -= iA * LA;
  sB.v = mulAdd(sB.v, mB, P);
  sB.w += iB * LB;
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD_COLD void solveWheel(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  const bool fixedRotation = (iA + iB == 0.0f);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 d = add(add(sub(sB.dp, sA.dp), j.deltaCenter), sub(rB, rA));
  const Vec2 axisA = rotate(sA.dq, j.axisA);
  const float translation = dot(axisA, d);
  const float a1 = cross(add(d, rA), axisA);
  const float a2 = cross(rB, axisA);
  if (j.enableMotor && fixedRotation == false) {
    const float Cdot = wB - wA - j.motorSpeed;
    float impulse = -j.motorMass * Cdot;
    const float oldImpulse = j.motorImpulse;
    const float maxImpulse = h * j.maxMotorTorque;
    j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
    impulse = j.motorImpulse - oldImpulse;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  if (j.enableSpring) {
    // a real spring: applied in the relax iterations too
    const float C = translation;
    const float bias = j.springSoftness.biasRate * C;
    const float massScale = j.springSoftness.massScale;
    const float impulseScale = j.springSoftness.impulseScale;
    const float Cdot = dot(axisA, sub(vB, vA)) + a2 * wB - a1 * wA;
    const float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.springImpulse;
    j.springImpulse += impulse;
    const Vec2 P = mulSV(impulse, axisA);
    const float LA = impulse * a1;
    const float LB = impulse * a2;
    vA = mulSub(vA, mA, P);
    wA -= iA * LA;
    vB = mulAdd(vB, mB, P);
    wB += iB * LB;
  }
  if (j.enableLimit) {
    {
      // lower limit
      const float C = translation - j.lowerAngle;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;  // speculative
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      const float Cdot = dot(axisA, sub(vB, vA)) + a2 * wB - a1 * wA;
      float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.lowerImpulse;
      const float oldImpulse = j.lowerImpulse;
      j.lowerImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.lowerImpulse - oldImpulse;
      const Vec2 P = mulSV(impulse, axisA);
      const float LA = impulse * a1;
      const float LB = impulse * a2;
      vA = mulSub(vA, mA, P);
      wA -= iA * LA;
      vB = mulAdd(vB, mB, P);
      wB += iB * LB;
    }
    {
      // upper limit, signs flipped so that C and the impulse stay positive
      const float C = j.upperAngle - translation;
      float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
      if (C > 0.0f) {
        bias = C * inv_h;
      } else if (useBias) {
        bias = j.constraintSoftness.biasRate * C;
        massScale = j.constraintSoftness.massScale;
        impulseScale = j.constraintSoftness.impulseScale;
      }
      const float Cdot = dot(axisA, sub(vA, vB)) + a1 * wA - a2 * wB;
      float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.upperImpulse;
      const float oldImpulse = j.upperImpulse;
      j.upperImpulse = maxf(oldImpulse + impulse, 0.0f);
      impulse = j.upperImpulse - oldImpulse;
      const Vec2 P = mulSV(impulse, axisA);
      const float LA = impulse * a1;
      const float LB = impulse * a2;
      vA = mulAdd(vA, mA, P);
      wA += iA * LA;
      vB = mulSub(vB, mB, P);
      wB -= iB * LB;
    }
  }
  {
    // point to line
    const Vec2 perpA = leftPerp(axisA);
    float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
    if (useBias) {
      const float C = dot(perpA, d);
      bias = j.constraintSoftness.biasRate * C;
      massScale = j.constraintSoftness.massScale;
      impulseScale = j.constraintSoftness.impulseScale;
    }
    const float s1 = cross(add(d, rA), perpA);
    const float s2 = cross(rB, perpA);
    const float Cdot = dot(perpA, sub(vB, vA)) + s2 * wB - s1 * wA;
    const float impulse = -massScale * j.perpMass * (Cdot + bias) - impulseScale * j.linearImpulse.x;
    j.linearImpulse.x += impulse;
    const Vec2 P = mulSV(impulse, perpA);
    const float LA = impulse * s1;
    const float LB = impulse * s2;
    vA = mulSub(vA, mA, P);
    wA -= iA * LA;
    vB = mulAdd(vB, mB, P);
    wB += iB * LB;
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// Distance joint stages, S/distance_joint.c: prepare :215-278, warm start :280-310,
// solve :312-490.
F2D_HD_COLD void prepareDistance(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  const float mA = a.invMass, iA = a.invI, mB = b.invMass, iB = b.invI;
  j.invMassA = mA;
  j.invMassB = mB;
  j.invIA = iA;
  j.invIB = iB;
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  j.anchorA = rotate(a.xf.q, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(b.xf.q, sub(j.localOriginAnchorB, b.localCenter));
  j.deltaCenter = sub(b.center, a.center);
  const Vec2 rA = j.anchorA, rB = j.anchorB;
  const Vec2 separation = add(sub(rB, rA), j.deltaCenter);
  const Vec2 axis = normalize(separation);
  const float crA = cross(rA, axis);
  const float crB = cross(rB, axis);
  const float k = mA + mB + iA * crA * crA + iB * crB * crB;
  j.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
  j.springSoftness = makeSoft(j.hertz, j.dampingRatio, h);
  if (warmStartScale == 0.0f) {
    j.linearImpulse.x = 0.0f;
    j.lowerImpulse = 0.0f;
    j.upperImpulse = 0.0f;
    j.motorImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartDistance(Joint& j, BodyState* states) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 ds = add(sub(sB.dp, sA.dp), sub(rB, rA));
  const Vec2 separation = add(j.deltaCenter, ds);
  const Vec2 axis = normalize(separation);
  const float axialImpulse = j.linearImpulse.x + j.lowerImpulse - j.upperImpulse + j.motorImpulse;
  const Vec2 P = mulSV(axialImpulse, axis);
  sA.v = mulSub(sA.v, mA, P);
  sA.w -= iA * cross(rA, P);
  sB.v = mulAdd(sB.v, mB, P);
  sB.w += iB * cross(rB, P);
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD_COLD void solveDistance(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v;
  float wA = sA.w;
  Vec2 vB = sB.v;
  float wB = sB.w;
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  const Vec2 ds = add(sub(sB.dp, sA.dp), sub(rB, rA));
  const Vec2 separation = add(j.deltaCenter, ds);
  const float len = length(separation);
  const Vec2 axis = normalize(separation);
  const float minLength = j.lowerAngle, maxLength = j.upperAngle;
  auto apply = [&](float impulse) {
    const Vec2 P = mulSV(impulse, axis);
    vA = mulSub(vA, mA, P);
    wA -= iA * cross(rA, P);
    vB = mulAdd(vB, mB, P);
    wB += iB * cross(rB, P);
  };
  if (j.enableSpring && (minLength < maxLength || j.enableLimit == false)) {
    // spring
    if (j.hertz > 0.0f) {
      const Vec2 vr = add(sub(vB, vA), sub(crossSV(wB, rB), crossSV(wA, rA)));
      const float Cdot = dot(axis, vr);
      const float C = len - j.targetAngle;
      const float bias = j.springSoftness.biasRate * C;
      const float m = j.springSoftness.massScale * j.axialMass;
      const float impulse = -m * (Cdot + bias) - j.springSoftness.impulseScale * j.linearImpulse.x;
      j.linearImpulse.x += impulse;
      apply(impulse);
    }
    if (j.enableLimit) {
      {
        // lower
        const Vec2 vr = add(sub(vB, vA), sub(crossSV(wB, rB), crossSV(wA, rA)));
        const float Cdot = dot(axis, vr);
        const float C = len - minLength;
        float bias = 0.0f, massCoeff = 1.0f, impulseCoeff = 0.0f;
        if (C > 0.0f) {
          bias = C * inv_h;  // speculative
        } else if (useBias) {
          bias = j.constraintSoftness.biasRate * C;
          massCoeff = j.constraintSoftness.massScale;
          impulseCoeff = j.constraintSoftness.impulseScale;
        }
        float impulse = -massCoeff * j.axialMass * (Cdot + bias) - impulseCoeff * j.lowerImpulse;
        const float newImpulse = maxf(0.0f, j.lowerImpulse + impulse);
        impulse = newImpulse - j.lowerImpulse;
        j.lowerImpulse = newImpulse;
        apply(impulse);
      }
      {
        // upper
        const Vec2 vr = add(sub(vA, vB), sub(crossSV(wA, rA), crossSV(wB, rB)));
        const float Cdot = dot(axis, vr);
        const float C = maxLength - len;
        float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
        if (C > 0.0f) {
          bias = C * inv_h;
        } else if (useBias) {
          bias = j.constraintSoftness.biasRate * C;
          massScale = j.constraintSoftness.massScale;
          impulseScale = j.constraintSoftness.impulseScale;
        }
        float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.upperImpulse;
        const float newImpulse = maxf(0.0f, j.upperImpulse + impulse);
        impulse = newImpulse - j.upperImpulse;
        j.upperImpulse = newImpulse;
        apply(-impulse);
      }
    }
    if (j.enableMotor) {
      const Vec2 vr = add(sub(vB, vA), sub(crossSV(wB, rB), crossSV(wA, rA)));
      const float Cdot = dot(axis, vr);
      float impulse = j.axialMass * (j.motorSpeed - Cdot);
      const float oldImpulse = j.motorImpulse;
      const float maxImpulse = h * j.maxMotorTorque;
      j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
      impulse = j.motorImpulse - oldImpulse;
      apply(impulse);
    }
  } else {
    // rigid constraint
    const Vec2 vr = add(sub(vB, vA), sub(crossSV(wB, rB), crossSV(wA, rA)));
    const float Cdot = dot(axis, vr);
    const float C = len - j.targetAngle;
    float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
    if (useBias) {
      bias = j.constraintSoftness.biasRate * C;
      massScale = j.constraintSoftness.massScale;
      impulseScale = j.constraintSoftness.impulseScale;
    }
    const float impulse = -massScale * j.axialMass * (Cdot + bias) - impulseScale * j.linearImpulse.x;
    j.linearImpulse.x += impulse;
    apply(impulse);
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// b2GetInverse22 (I/math_functions.h:551-564) of K = [[k11, k12], [k12, k22]], written
// into the two columns (cx, cy) that b2MulMV reads (I/math_functions.h:541-548)
F2D_HD void inverse22(float k11, float k12, float k22, Vec2& cx, Vec2& cy) {
  const float a = k11, b = k12, c = k12, d = k22;
  float det = a * d - b * c;
  if (det != 0.0f) det = 1.0f / det;
  cx = {det * d, -det * c};
  cy = {-det * b, det * a};
}
F2D_HD Vec2 mulMV(Vec2 cx, Vec2 cy, Vec2 v) { return {cx.x * v.x + cy.x * v.y, cx.y * v.x + cy.y * v.y}; }

// Motor joint, S/motor_joint.c:99-254. Fields: localAxisA = linear offset,
// referenceAngle = angular offset, maxMotorTorque = maximum force, angularHertz =
// maximum torque, angularDampingRatio = correction factor; prepared linearMass in
// (axisA, {perpMass, motorMass}) as columns, angularMass in axialMass; the angular
// impulse in springImpulse.
F2D_HD_COLD void prepareMotor(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& a = sims[j.bodyA];
  const BodySim& b = sims[j.bodyB];
  const float mA = a.invMass, iA = a.invI, mB = b.invMass, iB = b.invI;
  j.invMassA = mA;
  j.invMassB = mB;
  j.invIA = iA;
  j.invIB = iB;
  j.simA = a.setIndex == kAwakeSet ? j.bodyA : -1;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  j.anchorA = rotate(a.xf.q, sub(j.localOriginAnchorA, a.localCenter));
  j.anchorB = rotate(b.xf.q, sub(j.localOriginAnchorB, b.localCenter));
  j.deltaCenter = sub(sub(b.center, a.center), j.localAxisA);
  j.deltaAngle = relativeAngle(b.xf.q, a.xf.q) - j.referenceAngle;
  const Vec2 rA = j.anchorA, rB = j.anchorB;
  const float k11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
  const float k12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
  const float k22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
  Vec2 cy;
  inverse22(k11, k12, k22, j.axisA, cy);
  j.perpMass = cy.x;
  j.motorMass = cy.y;
  const float ka = iA + iB;
  j.axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
  if (warmStartScale == 0.0f) {
    j.linearImpulse = {0.0f, 0.0f};
    j.springImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartMotor(Joint& j, BodyState* states) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  BodyState sA = loadState(states, ia);
  BodyState sB = loadState(states, ib);
  const Vec2 rA = rotate(sA.dq, j.anchorA);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  sA.v = mulSub(sA.v, mA, j.linearImpulse);
  sA.w -= iA * (cross(rA, j.linearImpulse) + j.springImpulse);
  sB.v = mulAdd(sB.v, mB, j.linearImpulse);
  sB.w += iB * (cross(rB, j.linearImpulse) + j.springImpulse);
  storeVel(states, ia, sA.v, sA.w);
  storeVel(states, ib, sB.v, sB.w);
}

F2D_HD_COLD void solveMotor(Joint& j, BodyState* states, float h, float inv_h) {
  const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
  const int ia = j.simA, ib = j.simB;
  const BodyState sA = loadState(states, ia);
  const BodyState sB = loadState(states, ib);
  Vec2 vA = sA.v, vB = sB.v;
  float wA = sA.w, wB = sB.w;
  const float correction = j.angularDampingRatio;
  {
    float angularSeparation = relativeAngle(sB.dq, sA.dq) + j.deltaAngle;
    angularSeparation = unwindAngle(angularSeparation);
    const float angularBias = inv_h * correction * angularSeparation;
    const float Cdot = wB - wA;
    float impulse = -j.axialMass * (Cdot + angularBias);
    const float oldImpulse = j.springImpulse;
    const float maxImpulse = h * j.angularHertz;
    j.springImpulse = clampf(j.springImpulse + impulse, -maxImpulse, maxImpulse);
    impulse = j.springImpulse - oldImpulse;
    wA -= iA * impulse;
    wB += iB * impulse;
  }
  {
    const Vec2 rA = rotate(sA.dq, j.anchorA);
    const Vec2 rB = rotate(sB.dq, j.anchorB);
    const Vec2 ds = add(sub(sB.dp, sA.dp), sub(rB, rA));
    const Vec2 linearSeparation = add(j.deltaCenter, ds);
    const Vec2 linearBias = mulSV(inv_h * correction, linearSeparation);
    const Vec2 Cdot = sub(add(vB, crossSV(wB, rB)), add(vA, crossSV(wA, rA)));
    const Vec2 b = mulMV(j.axisA, {j.perpMass, j.motorMass}, add(Cdot, linearBias));
    Vec2 impulse = {-b.x, -b.y};
    const Vec2 oldImpulse = j.linearImpulse;
    const float maxImpulse = h * j.maxMotorTorque;
    j.linearImpulse = add(j.linearImpulse, impulse);
    if (lengthSq(j.linearImpulse) > maxImpulse * maxImpulse) {
      j.linearImpulse = normalize(j.linearImpulse);
      j.linearImpulse.x *= maxImpulse;
      j.linearImpulse.y *= maxImpulse;
    }
    impulse = sub(j.linearImpulse, oldImpulse);
    vA = mulSub(vA, mA, impulse);
    wA -= iA * cross(rA, impulse);
    vB = mulAdd(vB, mB, impulse);
    wB += iB * cross(rB, impulse);
  }
  storeVel(states, ia, vA, wA);
  storeVel(states, ib, vB, wB);
}

// Mouse joint, S/mouse_joint.c:77-214: body B only. Fields: localAxisA = target,
// maxMotorTorque = maximum force, springSoftness = linear softness, angularSoftness
// (0.5 Hz, damping 0.1); prepared linearMass in (axisA, {perpMass, motorMass}); the
// angular impulse in springImpulse.
F2D_HD_COLD void prepareMouse(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
  const BodySim& b = sims[j.bodyB];
  j.invMassB = b.invMass;
  j.invIB = b.invI;
  j.simB = b.setIndex == kAwakeSet ? j.bodyB : -1;
  j.anchorB = rotate(b.xf.q, sub(j.localOriginAnchorB, b.localCenter));
  j.springSoftness = makeSoft(j.hertz, j.dampingRatio, h);
  j.angularSoftness = makeSoft(0.5f, 0.1f, h);
  const Vec2 rB = j.anchorB;
  const float mB = b.invMass, iB = b.invI;
  const float k11 = mB + iB * rB.y * rB.y;
  const float k12 = -iB * rB.x * rB.y;
  const float k22 = mB + iB * rB.x * rB.x;
  Vec2 cy;
  inverse22(k11, k12, k22, j.axisA, cy);
  j.perpMass = cy.x;
  j.motorMass = cy.y;
  j.deltaCenter = sub(b.center, j.localAxisA);
  if (warmStartScale == 0.0f) {
    j.linearImpulse = {0.0f, 0.0f};
    j.springImpulse = 0.0f;
  }
}

F2D_HD_COLD void warmStartMouse(Joint& j, BodyState* states) {
  const float mB = j.invMassB, iB = j.invIB;
  const int ib = j.simB;
  BodyState sB = loadState(states, ib);
  const Vec2 rB = rotate(sB.dq, j.anchorB);
  Vec2 vB = mulAdd(sB.v, mB, j.linearImpulse);
  float wB = sB.w + iB * (cross(rB, j.linearImpulse) + j.springImpulse);
  storeVel(states, ib, vB, wB);
}

F2D_HD_COLD void solveMouse(Joint& j, BodyState* states, float h) {
  const float mB = j.invMassB, iB = j.invIB;
  const int ib = j.simB;
  const BodyState sB = loadState(states, ib);
  Vec2 vB = sB.v;
  float wB = sB.w;
  {
    const float massScale = j.angularSoftness.massScale;
    const float impulseScale = j.angularSoftness.impulseScale;
    float impulse = iB > 0.0f ? -wB / iB : 0.0f;
    impulse = massScale * impulse - impulseScale * j.springImpulse;
    j.springImpulse += impulse;
    wB += iB * impulse;
  }
  const float maxImpulse = j.maxMotorTorque * h;
  {
    const Vec2 rB = rotate(sB.dq, j.anchorB);
    const Vec2 Cdot = add(vB, crossSV(wB, rB));
    const Vec2 separation = add(add(sB.dp, rB), j.deltaCenter);
    const Vec2 bias = mulSV(j.springSoftness.biasRate, separation);
    const float massScale = j.springSoftness.massScale;
    const float impulseScale = j.springSoftness.impulseScale;
    const Vec2 b = mulMV(j.axisA, {j.perpMass, j.motorMass}, add(Cdot, bias));
    Vec2 impulse;
    impulse.x = -massScale * b.x - impulseScale * j.linearImpulse.x;
    impulse.y = -massScale * b.y - impulseScale * j.linearImpulse.y;
    const Vec2 oldImpulse = j.linearImpulse;
    j.linearImpulse.x += impulse.x;
    j.linearImpulse.y += impulse.y;
    const float mag = length(j.linearImpulse);
    if (mag > maxImpulse) j.linearImpulse = mulSV(maxImpulse, normalize(j.linearImpulse));
    impulse.x = j.linearImpulse.x - oldImpulse.x;
    impulse.y = j.linearImpulse.y - oldImpulse.y;
    vB = mulAdd(vB, mB, impulse);
    wB += iB * cross(rB, impulse);
  }
  storeVel(states, ib, vB, wB);
}

// Filter joint: only the common constraint softness of b2PrepareJoint
// (S/joint.c:1348-1352); no warm start or solve.
F2D_HD_COLD void prepareFilter(Joint& j, float h, float inv_h) {
  float hertz = minf(j.constraintHertz, 0.25f * inv_h);
  j.constraintSoftness = makeSoft(hertz, j.constraintDampingRatio, h);
}

// The joint stages by joint type (b2PrepareJointsTask / b2WarmStartJoint /
// b2SolveJoint, S/joint.c:1348-1450). Revolute joints take an inlined path; the other
// types are reached through out-of-line calls, which keeps the solve kernel at the
// register count of the revolute code (inlined together, the five types raise it from
// 77 to 248 VGPRs and the occupancy from 6 to 2 waves per SIMD).
F2D_HD_COLD void prepareJointOther(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  switch (j.type) {
    case kPrismaticJoint: preparePrismatic(j, sims, h, inv_h, warmStartScale); break;
    case kWeldJoint: prepareWeld(j, sims, h, inv_h, warmStartScale); break;
    case kWheelJoint: prepareWheel(j, sims, h, inv_h, warmStartScale); break;
    case kDistanceJoint: prepareDistance(j, sims, h, inv_h, warmStartScale); break;
    case kMotorJoint: prepareMotor(j, sims, h, inv_h, warmStartScale); break;
    case kMouseJoint: prepareMouse(j, sims, h, inv_h, warmStartScale); break;
    case kFilterJoint: prepareFilter(j, h, inv_h); break;
    default: prepareRevolute(j, sims, h, inv_h, warmStartScale); break;
  }
}
F2D_HD_COLD void warmStartJointOther(Joint& j, BodyState* states) {
  switch (j.type) {
    case kPrismaticJoint: warmStartPrismatic(j, states); break;
    case kWeldJoint: warmStartWeld(j, states); break;
    case kWheelJoint: warmStartWheel(j, states); break;
    case kDistanceJoint: warmStartDistance(j, states); break;
    case kMotorJoint: warmStartMotor(j, states); break;
    case kMouseJoint: warmStartMouse(j, states); break;
    case kFilterJoint: break;
    default: warmStartRevolute(j, states); break;
  }
}
F2D_HD_COLD void solveJointOther(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  switch (j.type) {
    case kPrismaticJoint: solvePrismatic(j, states, h, inv_h, useBias); break;
    case kWeldJoint: solveWeld(j, states, useBias); break;
    case kWheelJoint: solveWheel(j, states, h, inv_h, useBias); break;
    case kDistanceJoint: solveDistance(j, states, h, inv_h, useBias); break;
    case kMotorJoint: solveMotor(j, states, h, inv_h); break;
    case kMouseJoint: solveMouse(j, states, h); break;
    case kFilterJoint: break;
    default: solveRevolute(j, states, h, inv_h, useBias); break;
  }
}
F2D_HD void prepareJoint(Joint& j, const BodySim* sims, float h, float inv_h, float warmStartScale) {
  if (j.type == kRevoluteJoint) prepareRevolute(j, sims, h, inv_h, warmStartScale);
  else prepareJointOther(j, sims, h, inv_h, warmStartScale);
}
F2D_HD void warmStartJoint(Joint& j, BodyState* states) {
  if (j.type == kRevoluteJoint) warmStartRevolute(j, states);
  else warmStartJointOther(j, states);
}
F2D_HD void solveJoint(Joint& j, BodyState* states, float h, float inv_h, bool useBias) {
  if (j.type == kRevoluteJoint) solveRevolute(j, states, h, inv_h, useBias);
  else solveJointOther(j, states, h, inv_h, useBias);
}

}  // namespace f2d
