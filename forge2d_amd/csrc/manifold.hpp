// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Per-pair contact manifolds with Box2D v3.1 feature ids.
//
// Each function computes the manifold of one shape pair with the operation
// order of its reference counterpart in S/manifold.c (cited per function), so
// point counts, feature ids and every float match the reference bit for bit.
// Feature ids: MAKE_ID(a, b) = a << 8 | b (S/manifold.c:14).
#pragma once

#include "layout.hpp"

namespace f2d {

F2D_HD int makeId(int a, int b) { return ((a & 0xFF) << 8) | (b & 0xFF); }

// Polygon loops run over the fixed maximum of 8 vertices with a guard on the
// live count, and runtime-indexed reads are select chains, so the vertex arrays
// stay in registers. Live iterations execute the same operations in the same
// order as the reference's loops.
#if defined(__HIP_DEVICE_COMPILE__)
#define F2D_UNROLL _Pragma("unroll")
#else
#define F2D_UNROLL
#endif

// Select a[i] for 0 <= i < 8 with a three-level tree of selects on the bits of i.
F2D_HD float pickf(float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7, int i) {
  float t0 = (i & 1) ? a1 : a0;
  float t1 = (i & 1) ? a3 : a2;
  float t2 = (i & 1) ? a5 : a4;
  float t3 = (i & 1) ? a7 : a6;
  float u0 = (i & 2) ? t1 : t0;
  float u1 = (i & 2) ? t3 : t2;
  return (i & 4) ? u1 : u0;
}
F2D_HD Vec2 pick(const Vec2* a, int i) {
  return {pickf(a[0].x, a[1].x, a[2].x, a[3].x, a[4].x, a[5].x, a[6].x, a[7].x, i),
          pickf(a[0].y, a[1].y, a[2].y, a[3].y, a[4].y, a[5].y, a[6].y, a[7].y, i)};
}

struct Poly {
  Vec2 v[kMaxPolygonVertices];
  Vec2 n[kMaxPolygonVertices];
  Vec2 centroid;
  float radius;
  int count;
};

F2D_HD void zeroManifold(Manifold& m) {
  m.normal = {0.0f, 0.0f};
  m.rollingImpulse = 0.0f;
  m.pointCount = 0;
  for (int i = 0; i < 2; ++i) {
    ManifoldPoint& p = m.points[i];
    p.point = {0.0f, 0.0f};
    p.anchorA = {0.0f, 0.0f};
    p.anchorB = {0.0f, 0.0f};
    p.separation = 0.0f;
    p.normalImpulse = 0.0f;
    p.tangentImpulse = 0.0f;
    p.totalNormalImpulse = 0.0f;
    p.normalVelocity = 0.0f;
    p.id = 0;
  }
}

// b2MakeCapsule, S/manifold.c:16-35
F2D_HD Poly makeCapsulePoly(Vec2 p1, Vec2 p2, float radius) {
  Poly s;
  s.v[0] = p1;
  s.v[1] = p2;
  s.centroid = lerp(p1, p2, 0.5f);
  Vec2 d = sub(p2, p1);
  Vec2 axis = normalize(d);
  Vec2 nrm = rightPerp(axis);
  s.n[0] = nrm;
  s.n[1] = neg(nrm);
  s.count = 2;
  s.radius = radius;
  return s;
}

// b2CollideCircles, S/manifold.c:40-74
F2D_HD void collideCircles(Manifold& m, Vec2 centerA, float radiusA, Xf xfA, Vec2 centerB, float radiusB,
                           Xf xfB) {
  zeroManifold(m);
  Xf xf = invMulTransforms(xfA, xfB);
  Vec2 pointA = centerA;
  Vec2 pointB = transformPoint(xf, centerB);
  float distance;
  Vec2 normal = lengthAndNormalize(&distance, sub(pointB, pointA));
  float separation = distance - radiusA - radiusB;
  if (separation > kSpeculativeDistance) return;
  Vec2 cA = mulAdd(pointA, radiusA, normal);
  Vec2 cB = mulAdd(pointB, -radiusB, normal);
  Vec2 contactPointA = lerp(cA, cB, 0.5f);
  m.normal = rotate(xfA.q, normal);
  ManifoldPoint& mp = m.points[0];
  mp.anchorA = rotate(xfA.q, contactPointA);
  mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
  mp.point = add(mp.anchorA, xfA.p);
  mp.separation = separation;
  mp.id = 0;
  m.pointCount = 1;
}

// b2CollideCapsuleAndCircle, S/manifold.c:77-139
F2D_HD void collideCapsuleAndCircle(Manifold& m, Vec2 p1, Vec2 p2, float radiusA, Xf xfA, Vec2 centerB,
                                    float radiusB, Xf xfB) {
  zeroManifold(m);
  Xf xf = invMulTransforms(xfA, xfB);
  Vec2 pB = transformPoint(xf, centerB);
  Vec2 e = sub(p2, p1);
  Vec2 pA;
  float s1 = dot(sub(pB, p1), e);
  float s2 = dot(sub(p2, pB), e);
  if (s1 < 0.0f) {
    pA = p1;
  } else if (s2 < 0.0f) {
    pA = p2;
  } else {
    float s = s1 / dot(e, e);
    pA = mulAdd(p1, s, e);
  }
  float distance;
  Vec2 normal = lengthAndNormalize(&distance, sub(pB, pA));
  float separation = distance - radiusA - radiusB;
  if (separation > kSpeculativeDistance) return;
  Vec2 cA = mulAdd(pA, radiusA, normal);
  Vec2 cB = mulAdd(pB, -radiusB, normal);
  Vec2 contactPointA = lerp(cA, cB, 0.5f);
  m.normal = rotate(xfA.q, normal);
  ManifoldPoint& mp = m.points[0];
  mp.anchorA = rotate(xfA.q, contactPointA);
  mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
  mp.point = add(xfA.p, mp.anchorA);
  mp.separation = separation;
  mp.id = 0;
  m.pointCount = 1;
}

// b2CollidePolygonAndCircle, S/manifold.c:141-259
F2D_HD void collidePolygonAndCircle(Manifold& m, const Poly& A, Xf xfA, Vec2 centerB, float radiusB, Xf xfB) {
  zeroManifold(m);
  const float speculativeDistance = kSpeculativeDistance;
  Xf xf = invMulTransforms(xfA, xfB);
  Vec2 center = transformPoint(xf, centerB);
  float radiusA = A.radius;
  float radius = radiusA + radiusB;
  int normalIndex = 0;
  float separation = -kFltMax;
  int vertexCount = A.count;
  F2D_UNROLL
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    if (i < vertexCount) {
      float s = dot(A.n[i], sub(center, A.v[i]));
      if (s > separation) {
        separation = s;
        normalIndex = i;
      }
    }
  }
  if (separation > radius + speculativeDistance) return;
  int vertIndex1 = normalIndex;
  int vertIndex2 = vertIndex1 + 1 < vertexCount ? vertIndex1 + 1 : 0;
  Vec2 v1 = pick(A.v, vertIndex1);
  Vec2 v2 = pick(A.v, vertIndex2);
  float u1 = dot(sub(center, v1), sub(v2, v1));
  float u2 = dot(sub(center, v2), sub(v1, v2));
  if (u1 < 0.0f && separation > kFltEpsilon) {
    Vec2 normal = normalize(sub(center, v1));
    separation = dot(sub(center, v1), normal);
    if (separation > radius + speculativeDistance) return;
    Vec2 cA = mulAdd(v1, radiusA, normal);
    Vec2 cB = mulSub(center, radiusB, normal);
    Vec2 contactPointA = lerp(cA, cB, 0.5f);
    m.normal = rotate(xfA.q, normal);
    ManifoldPoint& mp = m.points[0];
    mp.anchorA = rotate(xfA.q, contactPointA);
    mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
    mp.point = add(xfA.p, mp.anchorA);
    mp.separation = dot(sub(cB, cA), normal);
    mp.id = 0;
    m.pointCount = 1;
  } else if (u2 < 0.0f && separation > kFltEpsilon) {
    Vec2 normal = normalize(sub(center, v2));
    separation = dot(sub(center, v2), normal);
    if (separation > radius + speculativeDistance) return;
    Vec2 cA = mulAdd(v2, radiusA, normal);
    Vec2 cB = mulSub(center, radiusB, normal);
    Vec2 contactPointA = lerp(cA, cB, 0.5f);
    m.normal = rotate(xfA.q, normal);
    ManifoldPoint& mp = m.points[0];
    mp.anchorA = rotate(xfA.q, contactPointA);
    mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
    mp.point = add(xfA.p, mp.anchorA);
    mp.separation = dot(sub(cB, cA), normal);
    mp.id = 0;
    m.pointCount = 1;
  } else {
    Vec2 normal = pick(A.n, normalIndex);
    m.normal = rotate(xfA.q, normal);
    Vec2 cA = mulAdd(center, radiusA - dot(sub(center, v1), normal), normal);
    Vec2 cB = mulSub(center, radiusB, normal);
    Vec2 contactPointA = lerp(cA, cB, 0.5f);
    ManifoldPoint& mp = m.points[0];
    mp.anchorA = rotate(xfA.q, contactPointA);
    mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
    mp.point = add(xfA.p, mp.anchorA);
    mp.separation = separation - radius;
    mp.id = 0;
    m.pointCount = 1;
  }
}

// b2CollideCapsules, S/manifold.c:261-530 (Ericson 5.1.9 plus clipping)
F2D_HD void collideCapsules(Manifold& m, Vec2 a1, Vec2 a2, float radiusA, Xf xfA, Vec2 b1, Vec2 b2, float radiusB,
                            Xf xfB) {
  zeroManifold(m);
  Vec2 origin = a1;
  Xf sfA = {add(xfA.p, rotate(xfA.q, origin)), xfA.q};
  Xf xf = invMulTransforms(sfA, xfB);
  Vec2 p1 = {0.0f, 0.0f};
  Vec2 q1 = sub(a2, origin);
  Vec2 p2 = transformPoint(xf, b1);
  Vec2 q2 = transformPoint(xf, b2);
  Vec2 d1 = sub(q1, p1);
  Vec2 d2 = sub(q2, p2);
  float dd1 = dot(d1, d1);
  float dd2 = dot(d2, d2);
  const float epsSqr = kFltEpsilon * kFltEpsilon;
  Vec2 r = sub(p1, p2);
  float rd1 = dot(r, d1);
  float rd2 = dot(r, d2);
  float d12 = dot(d1, d2);
  float denom = dd1 * dd2 - d12 * d12;
  float f1 = 0.0f;
  if (denom != 0.0f) {
    f1 = clampf((d12 * rd2 - rd1 * dd2) / denom, 0.0f, 1.0f);
  }
  float f2 = (d12 * f1 + rd2) / dd2;
  if (f2 < 0.0f) {
    f2 = 0.0f;
    f1 = clampf(-rd1 / dd1, 0.0f, 1.0f);
  } else if (f2 > 1.0f) {
    f2 = 1.0f;
    f1 = clampf((d12 - rd1) / dd1, 0.0f, 1.0f);
  }
  Vec2 closest1 = mulAdd(p1, f1, d1);
  Vec2 closest2 = mulAdd(p2, f2, d2);
  float distanceSquared = distanceSq(closest1, closest2);
  float radius = radiusA + radiusB;
  float maxDistance = radius + kSpeculativeDistance;
  if (distanceSquared > maxDistance * maxDistance) return;
  float distance = sqrtf(distanceSquared);
  float length1, length2;
  Vec2 u1 = lengthAndNormalize(&length1, d1);
  Vec2 u2 = lengthAndNormalize(&length2, d2);
  float fp2 = dot(sub(p2, p1), u1);
  float fq2 = dot(sub(q2, p1), u1);
  bool outsideA = (fp2 <= 0.0f && fq2 <= 0.0f) || (fp2 >= length1 && fq2 >= length1);
  float fp1 = dot(sub(p1, p2), u2);
  float fq1 = dot(sub(q1, p2), u2);
  bool outsideB = (fp1 <= 0.0f && fq1 <= 0.0f) || (fp1 >= length2 && fq1 >= length2);
  if (outsideA == false && outsideB == false) {
    Vec2 normalA;
    float separationA;
    {
      normalA = leftPerp(u1);
      float ss1 = dot(sub(p2, p1), normalA);
      float ss2 = dot(sub(q2, p1), normalA);
      float s1p = ss1 < ss2 ? ss1 : ss2;
      float s1n = -ss1 < -ss2 ? -ss1 : -ss2;
      if (s1p > s1n) {
        separationA = s1p;
      } else {
        separationA = s1n;
        normalA = neg(normalA);
      }
    }
    Vec2 normalB;
    float separationB;
    {
      normalB = leftPerp(u2);
      float ss1 = dot(sub(p1, p2), normalB);
      float ss2 = dot(sub(q1, p2), normalB);
      float s1p = ss1 < ss2 ? ss1 : ss2;
      float s1n = -ss1 < -ss2 ? -ss1 : -ss2;
      if (s1p > s1n) {
        separationB = s1p;
      } else {
        separationB = s1n;
        normalB = neg(normalB);
      }
    }
    if (separationA + 0.1f * kLinearSlop >= separationB) {
      m.normal = normalA;
      Vec2 cp = p2;
      Vec2 cq = q2;
      if (fp2 < 0.0f && fq2 > 0.0f) {
        cp = lerp(p2, q2, (0.0f - fp2) / (fq2 - fp2));
      } else if (fq2 < 0.0f && fp2 > 0.0f) {
        cq = lerp(q2, p2, (0.0f - fq2) / (fp2 - fq2));
      }
      if (fp2 > length1 && fq2 < length1) {
        cp = lerp(p2, q2, (fp2 - length1) / (fp2 - fq2));
      } else if (fq2 > length1 && fp2 < length1) {
        cq = lerp(q2, p2, (fq2 - length1) / (fq2 - fp2));
      }
      float sp = dot(sub(cp, p1), normalA);
      float sq = dot(sub(cq, p1), normalA);
      if (sp <= distance + kLinearSlop || sq <= distance + kLinearSlop) {
        ManifoldPoint& mp0 = m.points[0];
        mp0.anchorA = mulAdd(cp, 0.5f * (radiusA - radiusB - sp), normalA);
        mp0.separation = sp - radius;
        mp0.id = makeId(0, 0);
        ManifoldPoint& mp1 = m.points[1];
        mp1.anchorA = mulAdd(cq, 0.5f * (radiusA - radiusB - sq), normalA);
        mp1.separation = sq - radius;
        mp1.id = makeId(0, 1);
        m.pointCount = 2;
      }
    } else {
      m.normal = neg(normalB);
      Vec2 cp = p1;
      Vec2 cq = q1;
      if (fp1 < 0.0f && fq1 > 0.0f) {
        cp = lerp(p1, q1, (0.0f - fp1) / (fq1 - fp1));
      } else if (fq1 < 0.0f && fp1 > 0.0f) {
        cq = lerp(q1, p1, (0.0f - fq1) / (fp1 - fq1));
      }
      if (fp1 > length2 && fq1 < length2) {
        cp = lerp(p1, q1, (fp1 - length2) / (fp1 - fq1));
      } else if (fq1 > length2 && fp1 < length2) {
        cq = lerp(q1, p1, (fq1 - length2) / (fq1 - fp1));
      }
      float sp = dot(sub(cp, p2), normalB);
      float sq = dot(sub(cq, p2), normalB);
      if (sp <= distance + kLinearSlop || sq <= distance + kLinearSlop) {
        ManifoldPoint& mp0 = m.points[0];
        mp0.anchorA = mulAdd(cp, 0.5f * (radiusB - radiusA - sp), normalB);
        mp0.separation = sp - radius;
        mp0.id = makeId(0, 0);
        ManifoldPoint& mp1 = m.points[1];
        mp1.anchorA = mulAdd(cq, 0.5f * (radiusB - radiusA - sq), normalB);
        mp1.separation = sq - radius;
        mp1.id = makeId(1, 0);
        m.pointCount = 2;
      }
    }
  }
  if (m.pointCount == 0) {
    Vec2 normal = sub(closest2, closest1);
    if (dot(normal, normal) > epsSqr) {
      normal = normalize(normal);
    } else {
      normal = leftPerp(u1);
    }
    Vec2 c1 = mulAdd(closest1, radiusA, normal);
    Vec2 c2 = mulAdd(closest2, -radiusB, normal);
    int i1 = f1 == 0.0f ? 0 : 1;
    int i2 = f2 == 0.0f ? 0 : 1;
    m.normal = normal;
    m.points[0].anchorA = lerp(c1, c2, 0.5f);
    m.points[0].separation = sqrtf(distanceSquared) - radius;
    m.points[0].id = makeId(i1, i2);
    m.pointCount = 1;
  }
  m.normal = rotate(xfA.q, m.normal);
  F2D_UNROLL
  for (int i = 0; i < 2; ++i) {
    if (i < m.pointCount) {
      ManifoldPoint& mp = m.points[i];
      mp.anchorA = rotate(xfA.q, add(mp.anchorA, origin));
      mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
      mp.point = add(xfA.p, mp.anchorA);
    }
  }
}

// b2SegmentDistance, S/distance.c:33-106
struct SegmentDistanceResult {
  Vec2 closest1, closest2;
  float fraction1, fraction2, distanceSquared;
};

F2D_HD SegmentDistanceResult segmentDistance(Vec2 p1, Vec2 q1, Vec2 p2, Vec2 q2) {
  SegmentDistanceResult result = {{0.0f, 0.0f}, {0.0f, 0.0f}, 0.0f, 0.0f, 0.0f};
  Vec2 d1 = sub(q1, p1);
  Vec2 d2 = sub(q2, p2);
  Vec2 r = sub(p1, p2);
  float dd1 = dot(d1, d1);
  float dd2 = dot(d2, d2);
  float rd1 = dot(r, d1);
  float rd2 = dot(r, d2);
  const float epsSqr = kFltEpsilon * kFltEpsilon;
  if (dd1 < epsSqr || dd2 < epsSqr) {
    if (dd1 >= epsSqr) {
      result.fraction1 = clampf(-rd1 / dd1, 0.0f, 1.0f);
      result.fraction2 = 0.0f;
    } else if (dd2 >= epsSqr) {
      result.fraction1 = 0.0f;
      result.fraction2 = clampf(rd2 / dd2, 0.0f, 1.0f);
    } else {
      result.fraction1 = 0.0f;
      result.fraction2 = 0.0f;
    }
  } else {
    float d12 = dot(d1, d2);
    float denominator = dd1 * dd2 - d12 * d12;
    float f1 = 0.0f;
    if (denominator != 0.0f) {
      f1 = clampf((d12 * rd2 - rd1 * dd2) / denominator, 0.0f, 1.0f);
    }
    float f2 = (d12 * f1 + rd2) / dd2;
    if (f2 < 0.0f) {
      f2 = 0.0f;
      f1 = clampf(-rd1 / dd1, 0.0f, 1.0f);
    } else if (f2 > 1.0f) {
      f2 = 1.0f;
      f1 = clampf((d12 - rd1) / dd1, 0.0f, 1.0f);
    }
    result.fraction1 = f1;
    result.fraction2 = f2;
  }
  result.closest1 = mulAdd(p1, result.fraction1, d1);
  result.closest2 = mulAdd(p2, result.fraction2, d2);
  result.distanceSquared = distanceSq(result.closest1, result.closest2);
  return result;
}

// b2ClipPolygons, S/manifold.c:545-678
F2D_HD void clipPolygons(Manifold& m, const Poly& polyA, const Poly& polyB, int edgeA, int edgeB, bool flip) {
  zeroManifold(m);
  int i11, i12, i21, i22;
  Vec2 normal, v11, v12, v21, v22;
  float r1, r2;
  if (flip) {
    i11 = edgeB;
    i12 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
    i21 = edgeA;
    i22 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
    normal = pick(polyB.n, i11);
    v11 = pick(polyB.v, i11);
    v12 = pick(polyB.v, i12);
    v21 = pick(polyA.v, i21);
    v22 = pick(polyA.v, i22);
    r1 = polyB.radius;
    r2 = polyA.radius;
  } else {
    i11 = edgeA;
    i12 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
    i21 = edgeB;
    i22 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
    normal = pick(polyA.n, i11);
    v11 = pick(polyA.v, i11);
    v12 = pick(polyA.v, i12);
    v21 = pick(polyB.v, i21);
    v22 = pick(polyB.v, i22);
    r1 = polyA.radius;
    r2 = polyB.radius;
  }
  Vec2 tangent = crossSV(1.0f, normal);
  float lower1 = 0.0f;
  float upper1 = dot(sub(v12, v11), tangent);
  float upper2 = dot(sub(v21, v11), tangent);
  float lower2 = dot(sub(v22, v11), tangent);
  if (upper2 < lower1 || upper1 < lower2) return;
  Vec2 vLower;
  if (lower2 < lower1 && upper2 - lower2 > kFltEpsilon) {
    vLower = lerp(v22, v21, (lower1 - lower2) / (upper2 - lower2));
  } else {
    vLower = v22;
  }
  Vec2 vUpper;
  if (upper2 > upper1 && upper2 - lower2 > kFltEpsilon) {
    vUpper = lerp(v22, v21, (upper1 - lower2) / (upper2 - lower2));
  } else {
    vUpper = v21;
  }
  float separationLower = dot(sub(vLower, v11), normal);
  float separationUpper = dot(sub(vUpper, v11), normal);
  vLower = mulAdd(vLower, 0.5f * (r1 - r2 - separationLower), normal);
  vUpper = mulAdd(vUpper, 0.5f * (r1 - r2 - separationUpper), normal);
  float radius = r1 + r2;
  if (flip == false) {
    m.normal = normal;
    m.points[0].anchorA = vLower;
    m.points[0].separation = separationLower - radius;
    m.points[0].id = makeId(i11, i22);
    m.points[1].anchorA = vUpper;
    m.points[1].separation = separationUpper - radius;
    m.points[1].id = makeId(i12, i21);
    m.pointCount = 2;
  } else {
    m.normal = neg(normal);
    m.points[0].anchorA = vUpper;
    m.points[0].separation = separationUpper - radius;
    m.points[0].id = makeId(i21, i12);
    m.points[1].anchorA = vLower;
    m.points[1].separation = separationLower - radius;
    m.points[1].id = makeId(i22, i11);
    m.pointCount = 2;
  }
}

// b2FindMaxSeparation, S/manifold.c:680-734
F2D_HD float findMaxSeparation(int* edgeIndex, const Poly& poly1, const Poly& poly2) {
  int count1 = poly1.count;
  int count2 = poly2.count;
  int bestIndex = 0;
  float maxSeparation = -kFltMax;
  F2D_UNROLL
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    if (i < count1) {
      Vec2 n = poly1.n[i];
      Vec2 v1 = poly1.v[i];
      float si = kFltMax;
      F2D_UNROLL
      for (int j = 0; j < kMaxPolygonVertices; ++j) {
        if (j < count2) {
          float sij = dot(n, sub(poly2.v[j], v1));
          if (sij < si) si = sij;
        }
      }
      if (si > maxSeparation) {
        maxSeparation = si;
        bestIndex = i;
      }
    }
  }
  *edgeIndex = bestIndex;
  return maxSeparation;
}

F2D_HD void vertexVertex(Manifold& m, Vec2 va, Vec2 vb, float radiusA, float radiusB, float distance, float radius,
                         int id) {
  Vec2 normal = sub(vb, va);
  float invDistance = 1.0f / distance;
  normal.x *= invDistance;
  normal.y *= invDistance;
  Vec2 c1 = mulAdd(va, radiusA, normal);
  Vec2 c2 = mulAdd(vb, -radiusB, normal);
  m.normal = normal;
  m.points[0].anchorA = lerp(c1, c2, 0.5f);
  m.points[0].separation = distance - radius;
  m.points[0].id = id;
  m.pointCount = 1;
}

// b2CollidePolygons, S/manifold.c:736-1075 (the active #if 1 branch)
F2D_HD void collidePolygons(Manifold& m, const Poly& polygonA, Xf xfA, const Poly& polygonB, Xf xfB) {
  zeroManifold(m);
  Vec2 origin = polygonA.v[0];
  const float linearSlop = kLinearSlop;
  const float speculativeDistance = kSpeculativeDistance;
  Xf sfA = {add(xfA.p, rotate(xfA.q, origin)), xfA.q};
  Xf xf = invMulTransforms(sfA, xfB);
  Poly localPolyA;
  localPolyA.count = polygonA.count;
  localPolyA.radius = polygonA.radius;
  localPolyA.v[0] = {0.0f, 0.0f};
  localPolyA.n[0] = polygonA.n[0];
  F2D_UNROLL
  for (int i = 1; i < kMaxPolygonVertices; ++i) {
    localPolyA.v[i] = sub(polygonA.v[i], origin);
    localPolyA.n[i] = polygonA.n[i];
  }
  Poly localPolyB;
  localPolyB.count = polygonB.count;
  localPolyB.radius = polygonB.radius;
  F2D_UNROLL
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    localPolyB.v[i] = transformPoint(xf, polygonB.v[i]);
    localPolyB.n[i] = rotate(xf.q, polygonB.n[i]);
  }
  int edgeA = 0;
  float separationA = findMaxSeparation(&edgeA, localPolyA, localPolyB);
  int edgeB = 0;
  float separationB = findMaxSeparation(&edgeB, localPolyB, localPolyA);
  float radius = localPolyA.radius + localPolyB.radius;
  if (separationA > speculativeDistance + radius || separationB > speculativeDistance + radius) return;
  bool flip;
  if (separationA >= separationB) {
    flip = false;
    Vec2 searchDirection = pick(localPolyA.n, edgeA);
    int count = localPolyB.count;
    edgeB = 0;
    float minDot = kFltMax;
    F2D_UNROLL
    for (int i = 0; i < kMaxPolygonVertices; ++i) {
      if (i < count) {
        float d = dot(searchDirection, localPolyB.n[i]);
        if (d < minDot) {
          minDot = d;
          edgeB = i;
        }
      }
    }
  } else {
    flip = true;
    Vec2 searchDirection = pick(localPolyB.n, edgeB);
    int count = localPolyA.count;
    edgeA = 0;
    float minDot = kFltMax;
    F2D_UNROLL
    for (int i = 0; i < kMaxPolygonVertices; ++i) {
      if (i < count) {
        float d = dot(searchDirection, localPolyA.n[i]);
        if (d < minDot) {
          minDot = d;
          edgeA = i;
        }
      }
    }
  }
  if (separationA > 0.1f * linearSlop || separationB > 0.1f * linearSlop) {
    int i11 = edgeA;
    int i12 = edgeA + 1 < localPolyA.count ? edgeA + 1 : 0;
    int i21 = edgeB;
    int i22 = edgeB + 1 < localPolyB.count ? edgeB + 1 : 0;
    Vec2 v11 = pick(localPolyA.v, i11);
    Vec2 v12 = pick(localPolyA.v, i12);
    Vec2 v21 = pick(localPolyB.v, i21);
    Vec2 v22 = pick(localPolyB.v, i22);
    SegmentDistanceResult result = segmentDistance(v11, v12, v21, v22);
    float distance = sqrtf(result.distanceSquared);
    float separation = distance - radius;
    if (distance - radius > speculativeDistance) return;
    clipPolygons(m, localPolyA, localPolyB, edgeA, edgeB, flip);
    float minSeparation = kFltMax;
    if (m.pointCount > 0) minSeparation = minf(minSeparation, m.points[0].separation);
    if (m.pointCount > 1) minSeparation = minf(minSeparation, m.points[1].separation);
    if (separation + 0.1f * linearSlop < minSeparation) {
      if (result.fraction1 == 0.0f && result.fraction2 == 0.0f) {
        vertexVertex(m, v11, v21, localPolyA.radius, localPolyB.radius, distance, radius, makeId(i11, i21));
      } else if (result.fraction1 == 0.0f && result.fraction2 == 1.0f) {
        vertexVertex(m, v11, v22, localPolyA.radius, localPolyB.radius, distance, radius, makeId(i11, i22));
      } else if (result.fraction1 == 1.0f && result.fraction2 == 0.0f) {
        vertexVertex(m, v12, v21, localPolyA.radius, localPolyB.radius, distance, radius, makeId(i12, i21));
      } else if (result.fraction1 == 1.0f && result.fraction2 == 1.0f) {
        vertexVertex(m, v12, v22, localPolyA.radius, localPolyB.radius, distance, radius, makeId(i12, i22));
      }
    }
  } else {
    clipPolygons(m, localPolyA, localPolyB, edgeA, edgeB, flip);
  }
  if (m.pointCount > 0) {
    m.normal = rotate(xfA.q, m.normal);
    F2D_UNROLL
    for (int i = 0; i < 2; ++i) {
      if (i < m.pointCount) {
        ManifoldPoint& mp = m.points[i];
        mp.anchorA = rotate(xfA.q, add(mp.anchorA, origin));
        mp.anchorB = add(mp.anchorA, sub(xfA.p, xfB.p));
        mp.point = add(xfA.p, mp.anchorA);
      }
    }
  }
}

F2D_HD void loadPoly(Poly& p, const ShapeGeom& g) {
  p.count = g.count;
  p.radius = g.radius;
  p.centroid = g.centroid;
  F2D_UNROLL
  for (int i = 0; i < kMaxPolygonVertices; ++i) {
    p.v[i] = g.v[i];
    p.n[i] = g.n[i];
  }
}

// Registered manifold function per (typeA, typeB), S/contact.c:151-184.
// Returns false for pairs with no manifold function (segment-segment).
F2D_HD bool hasManifoldFcn(int ta, int tb) { return !(ta == kSegment && tb == kSegment); }
// primary == the registration order, so a non-primary pair is flipped on creation
F2D_HD bool isPrimary(int ta, int tb) {
  if (ta == tb) return true;
  if (ta == kSegment) return true;
  if (tb == kSegment) return false;
  return ta > tb;
}

// Dispatch, S/contact.c:67-149. Shapes are already in registered (primary) order.
F2D_HD void computeManifold(Manifold& m, const ShapeGeom& gA, Xf xfA, const ShapeGeom& gB, Xf xfB) {
  int ta = gA.type, tb = gB.type;
  if (ta == kCircle) {  // circle-circle
    collideCircles(m, gA.v[0], gA.radius, xfA, gB.v[0], gB.radius, xfB);
  } else if (ta == kCapsule) {
    if (tb == kCircle) {
      collideCapsuleAndCircle(m, gA.v[0], gA.v[1], gA.radius, xfA, gB.v[0], gB.radius, xfB);
    } else {
      collideCapsules(m, gA.v[0], gA.v[1], gA.radius, xfA, gB.v[0], gB.v[1], gB.radius, xfB);
    }
  } else if (ta == kPolygon) {
    Poly A;
    loadPoly(A, gA);
    if (tb == kCircle) {
      collidePolygonAndCircle(m, A, xfA, gB.v[0], gB.radius, xfB);
    } else if (tb == kCapsule) {
      Poly B = makeCapsulePoly(gB.v[0], gB.v[1], gB.radius);
      collidePolygons(m, A, xfA, B, xfB);
    } else {
      Poly B;
      loadPoly(B, gB);
      collidePolygons(m, A, xfA, B, xfB);
    }
  } else {  // segment A
    if (tb == kCircle) {
      collideCapsuleAndCircle(m, gA.v[0], gA.v[1], 0.0f, xfA, gB.v[0], gB.radius, xfB);
    } else if (tb == kCapsule) {
      collideCapsules(m, gA.v[0], gA.v[1], 0.0f, xfA, gB.v[0], gB.v[1], gB.radius, xfB);
    } else {
      Poly A = makeCapsulePoly(gA.v[0], gA.v[1], 0.0f);
      Poly B;
      loadPoly(B, gB);
      collidePolygons(m, A, xfA, B, xfB);
    }
  }
}

// Shape radius used by rolling resistance, S/shape.c (b2GetShapeRadius)
F2D_HD float shapeRadius(const ShapeGeom& g) {
  if (g.type == kSegment) return 0.0f;
  return g.radius;
}

// b2ComputeShapeAABB, S/shape.c:578-599 with S/geometry.c:401-453
F2D_HD AABB computeShapeAABB(const ShapeGeom& g, Xf xf) {
  if (g.type == kCircle) {
    Vec2 p = transformPoint(xf, g.v[0]);
    float r = g.radius;
    return {{p.x - r, p.y - r}, {p.x + r, p.y + r}};
  }
  if (g.type == kCapsule) {
    Vec2 v1 = transformPoint(xf, g.v[0]);
    Vec2 v2 = transformPoint(xf, g.v[1]);
    Vec2 r = {g.radius, g.radius};
    return {sub(vmin(v1, v2), r), add(vmax(v1, v2), r)};
  }
  if (g.type == kPolygon) {
    Vec2 lower = transformPoint(xf, g.v[0]);
    Vec2 upper = lower;
    F2D_UNROLL
    for (int i = 1; i < kMaxPolygonVertices; ++i) {
      if (i < g.count) {
        Vec2 v = transformPoint(xf, g.v[i]);
        lower = vmin(lower, v);
        upper = vmax(upper, v);
      }
    }
    Vec2 r = {g.radius, g.radius};
    return {sub(lower, r), add(upper, r)};
  }
  Vec2 v1 = transformPoint(xf, g.v[0]);
  Vec2 v2 = transformPoint(xf, g.v[1]);
  return {vmin(v1, v2), vmax(v1, v2)};
}

}  // namespace f2d
