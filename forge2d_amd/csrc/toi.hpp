// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Time of impact for continuous collision: GJK distance (S/distance.c:421-606)
// and conservative advancement over separating axes (S/distance.c:931-1414), in
// the reference's operation order so the fractions agree bit for bit.
#pragma once

#include "layout.hpp"

namespace f2d {

// b2ShapeProxy: up to 8 support points and a radius
struct Proxy {
  Vec2 p[kMaxPolygonVertices];
  int count;
  float radius;
};

// b2MakeShapeDistanceProxy, S/shape.c:922-943 (via b2MakeProxy, S/distance.c:108-119)
F2D_HD Proxy makeShapeProxy(const ShapeGeom& g) {
  Proxy x;
  if (g.type == kCircle) {
    x.count = 1;
    x.p[0] = g.v[0];
    x.radius = g.radius;
  } else if (g.type == kCapsule || g.type == kSegment) {
    x.count = 2;
    x.p[0] = g.v[0];
    x.p[1] = g.v[1];
    x.radius = g.type == kCapsule ? g.radius : 0.0f;
  } else {
    x.count = mini(g.count, kMaxPolygonVertices);
    for (int i = 0; i < x.count; ++i) x.p[i] = g.v[i];
    x.radius = g.radius;
  }
  return x;
}

// b2Sweep and b2MakeSweep, S/body.h:180-189
struct Sweep {
  Vec2 localCenter, c1, c2;
  Rot q1, q2;
};

F2D_HD Sweep makeSweep(const BodySim& s) { return {s.localCenter, s.center0, s.center, s.rotation0, s.xf.q}; }

// b2GetSweepTransform, S/distance.c:14-30
F2D_HD Xf sweepTransform(const Sweep& s, float t) {
  Xf xf;
  xf.p = add(mulSV(1.0f - t, s.c1), mulSV(t, s.c2));
  Rot q = {(1.0f - t) * s.q1.c + t * s.q2.c, (1.0f - t) * s.q1.s + t * s.q2.s};
  xf.q = normalizeRot(q);
  xf.p = sub(xf.p, rotate(xf.q, s.localCenter));
  return xf;
}

// b2NLerp, I/math_functions.h:390-402
F2D_HD Rot nlerpRot(Rot q1, Rot q2, float t) {
  float omt = 1.0f - t;
  Rot q = {omt * q1.c + t * q2.c, omt * q1.s + t * q2.s};
  return normalizeRot(q);
}

// b2FindSupport, S/distance.c:148-166
F2D_HD int findSupport(const Proxy& x, Vec2 d) {
  int best = 0;
  float bestValue = dot(x.p[0], d);
  for (int i = 1; i < x.count; ++i) {
    float value = dot(x.p[i], d);
    if (value > bestValue) {
      best = i;
      bestValue = value;
    }
  }
  return best;
}

struct SimplexVertex {
  Vec2 wA, wB, w;
  float a;
  int indexA, indexB;
};

struct Simplex {
  SimplexVertex v[3];
  int count;
};

struct SimplexCache {
  int count;
  int indexA[3], indexB[3];
};

// b2ComputeSimplexWitnessPoints, S/distance.c:217-245 (b2Weight2/b2Weight3 order)
F2D_HD void witnessPoints(Vec2* a, Vec2* b, const Simplex& s) {
  if (s.count == 1) {
    *a = s.v[0].wA;
    *b = s.v[0].wB;
  } else if (s.count == 2) {
    *a = {s.v[0].a * s.v[0].wA.x + s.v[1].a * s.v[1].wA.x, s.v[0].a * s.v[0].wA.y + s.v[1].a * s.v[1].wA.y};
    *b = {s.v[0].a * s.v[0].wB.x + s.v[1].a * s.v[1].wB.x, s.v[0].a * s.v[0].wB.y + s.v[1].a * s.v[1].wB.y};
  } else if (s.count == 3) {
    *a = {s.v[0].a * s.v[0].wA.x + s.v[1].a * s.v[1].wA.x + s.v[2].a * s.v[2].wA.x,
          s.v[0].a * s.v[0].wA.y + s.v[1].a * s.v[1].wA.y + s.v[2].a * s.v[2].wA.y};
    *b = *a;
  } else {
    *a = {0.0f, 0.0f};
    *b = {0.0f, 0.0f};
  }
}

// b2SolveSimplex2, S/distance.c:271-304: closest point on a segment to the origin
F2D_HD Vec2 solveSimplex2(Simplex& s) {
  Vec2 w1 = s.v[0].w;
  Vec2 w2 = s.v[1].w;
  Vec2 e12 = sub(w2, w1);
  float d12_2 = -dot(w1, e12);
  if (d12_2 <= 0.0f) {
    s.v[0].a = 1.0f;
    s.count = 1;
    return neg(w1);
  }
  float d12_1 = dot(w2, e12);
  if (d12_1 <= 0.0f) {
    s.v[1].a = 1.0f;
    s.count = 1;
    s.v[0] = s.v[1];
    return neg(w2);
  }
  float inv_d12 = 1.0f / (d12_1 + d12_2);
  s.v[0].a = d12_1 * inv_d12;
  s.v[1].a = d12_2 * inv_d12;
  s.count = 2;
  return crossSV(cross(add(w1, w2), e12), e12);
}

// b2SolveSimplex3, S/distance.c:306-419: Voronoi regions of a triangle
F2D_HD Vec2 solveSimplex3(Simplex& s) {
  Vec2 w1 = s.v[0].w;
  Vec2 w2 = s.v[1].w;
  Vec2 w3 = s.v[2].w;
  Vec2 e12 = sub(w2, w1);
  float w1e12 = dot(w1, e12);
  float w2e12 = dot(w2, e12);
  float d12_1 = w2e12;
  float d12_2 = -w1e12;
  Vec2 e13 = sub(w3, w1);
  float w1e13 = dot(w1, e13);
  float w3e13 = dot(w3, e13);
  float d13_1 = w3e13;
  float d13_2 = -w1e13;
  Vec2 e23 = sub(w3, w2);
  float w2e23 = dot(w2, e23);
  float w3e23 = dot(w3, e23);
  float d23_1 = w3e23;
  float d23_2 = -w2e23;
  float n123 = cross(e12, e13);
  float d123_1 = n123 * cross(w2, w3);
  float d123_2 = n123 * cross(w3, w1);
  float d123_3 = n123 * cross(w1, w2);
  if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
    s.v[0].a = 1.0f;
    s.count = 1;
    return neg(w1);
  }
  if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
    float inv_d12 = 1.0f / (d12_1 + d12_2);
    s.v[0].a = d12_1 * inv_d12;
    s.v[1].a = d12_2 * inv_d12;
    s.count = 2;
    return crossSV(cross(add(w1, w2), e12), e12);
  }
  if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
    float inv_d13 = 1.0f / (d13_1 + d13_2);
    s.v[0].a = d13_1 * inv_d13;
    s.v[2].a = d13_2 * inv_d13;
    s.count = 2;
    s.v[1] = s.v[2];
    return crossSV(cross(add(w1, w3), e13), e13);
  }
  if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
    s.v[1].a = 1.0f;
    s.count = 1;
    s.v[0] = s.v[1];
    return neg(w2);
  }
  if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
    s.v[2].a = 1.0f;
    s.count = 1;
    s.v[0] = s.v[2];
    return neg(w3);
  }
  if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
    float inv_d23 = 1.0f / (d23_1 + d23_2);
    s.v[1].a = d23_1 * inv_d23;
    s.v[2].a = d23_2 * inv_d23;
    s.count = 2;
    s.v[0] = s.v[2];
    return crossSV(cross(add(w2, w3), e23), e23);
  }
  float inv_d123 = 1.0f / (d123_1 + d123_2 + d123_3);
  s.v[0].a = d123_1 * inv_d123;
  s.v[1].a = d123_2 * inv_d123;
  s.v[2].a = d123_3 * inv_d123;
  s.count = 3;
  return {0.0f, 0.0f};
}

// b2ShapeDistance with useRadii = false, S/distance.c:421-606. Returns the distance
// between the cores of the two proxies (0 when they overlap) and updates the cache
// exactly where the reference does (not on the early overlap exits).
F2D_HD float shapeDistance(SimplexCache& cache, const Proxy& proxyA, const Proxy& proxyB, Xf xfA, Xf xfB) {
  Proxy localB;
  {
    Xf t = invMulTransforms(xfA, xfB);
    localB.count = proxyB.count;
    localB.radius = proxyB.radius;
    for (int i = 0; i < localB.count; ++i) localB.p[i] = transformPoint(t, proxyB.p[i]);
  }
  // b2MakeSimplexFromCache, S/distance.c:168-204
  Simplex s;
  s.count = cache.count;
  for (int i = 0; i < s.count; ++i) {
    SimplexVertex& v = s.v[i];
    v.indexA = cache.indexA[i];
    v.indexB = cache.indexB[i];
    v.wA = proxyA.p[v.indexA];
    v.wB = localB.p[v.indexB];
    v.w = sub(v.wA, v.wB);
    v.a = -1.0f;
  }
  if (s.count == 0) {
    SimplexVertex& v = s.v[0];
    v.indexA = 0;
    v.indexB = 0;
    v.wA = proxyA.p[0];
    v.wB = localB.p[0];
    v.w = sub(v.wA, v.wB);
    v.a = 1.0f;
    s.count = 1;
  }
  Vec2 nonUnitNormal = {0.0f, 0.0f};
  int saveA[3], saveB[3];
  const int maxIterations = 20;
  int iteration = 0;
  while (iteration < maxIterations) {
    int saveCount = s.count;
    for (int i = 0; i < saveCount; ++i) {
      saveA[i] = s.v[i].indexA;
      saveB[i] = s.v[i].indexB;
    }
    Vec2 d = {0.0f, 0.0f};
    if (s.count == 1) {
      d = neg(s.v[0].w);
    } else if (s.count == 2) {
      d = solveSimplex2(s);
    } else {
      d = solveSimplex3(s);
    }
    if (s.count == 3) return 0.0f;  // origin inside the triangle: overlap
    if (dot(d, d) < kFltEpsilon * kFltEpsilon) return 0.0f;
    nonUnitNormal = d;
    SimplexVertex& v = s.v[s.count];
    v.indexA = findSupport(proxyA, d);
    v.wA = proxyA.p[v.indexA];
    v.indexB = findSupport(localB, neg(d));
    v.wB = localB.p[v.indexB];
    v.w = sub(v.wA, v.wB);
    ++iteration;
    bool duplicate = false;
    for (int i = 0; i < saveCount; ++i) {
      if (v.indexA == saveA[i] && v.indexB == saveB[i]) {
        duplicate = true;
        break;
      }
    }
    if (duplicate) break;
    s.count += 1;
  }
  (void)nonUnitNormal;
  Vec2 pa, pb;
  witnessPoints(&pa, &pb, s);
  float dx = pb.x - pa.x;
  float dy = pb.y - pa.y;
  float distance = sqrtf(dx * dx + dy * dy);  // b2Distance
  // b2MakeSimplexCache, S/distance.c:206-215
  cache.count = s.count;
  for (int i = 0; i < s.count; ++i) {
    cache.indexA[i] = s.v[i].indexA;
    cache.indexB[i] = s.v[i].indexB;
  }
  return distance;
}

enum SeparationType : int { kPointsType = 0, kFaceAType = 1, kFaceBType = 2 };

struct SeparationFunction {
  Vec2 localPoint, axis;
  int type;
};

// b2MakeSeparationFunction, S/distance.c:941-1016
F2D_HD SeparationFunction makeSeparation(const SimplexCache& cache, const Proxy& proxyA, const Sweep& sweepA,
                                         const Proxy& proxyB, const Sweep& sweepB, float t1) {
  SeparationFunction f;
  Xf xfA = sweepTransform(sweepA, t1);
  Xf xfB = sweepTransform(sweepB, t1);
  if (cache.count == 1) {
    f.type = kPointsType;
    Vec2 pointA = transformPoint(xfA, proxyA.p[cache.indexA[0]]);
    Vec2 pointB = transformPoint(xfB, proxyB.p[cache.indexB[0]]);
    f.axis = normalize(sub(pointB, pointA));
    f.localPoint = {0.0f, 0.0f};
    return f;
  }
  if (cache.indexA[0] == cache.indexA[1]) {
    // two points on B and one on A
    f.type = kFaceBType;
    Vec2 b1 = proxyB.p[cache.indexB[0]];
    Vec2 b2 = proxyB.p[cache.indexB[1]];
    f.axis = crossVS(sub(b2, b1), 1.0f);
    f.axis = normalize(f.axis);
    Vec2 normal = rotate(xfB.q, f.axis);
    f.localPoint = {0.5f * (b1.x + b2.x), 0.5f * (b1.y + b2.y)};
    Vec2 pointB = transformPoint(xfB, f.localPoint);
    Vec2 pointA = transformPoint(xfA, proxyA.p[cache.indexA[0]]);
    float sep = dot(sub(pointA, pointB), normal);
    if (sep < 0.0f) f.axis = neg(f.axis);
    return f;
  }
  // two points on A and one or two points on B
  f.type = kFaceAType;
  Vec2 a1 = proxyA.p[cache.indexA[0]];
  Vec2 a2 = proxyA.p[cache.indexA[1]];
  f.axis = crossVS(sub(a2, a1), 1.0f);
  f.axis = normalize(f.axis);
  Vec2 normal = rotate(xfA.q, f.axis);
  f.localPoint = {0.5f * (a1.x + a2.x), 0.5f * (a1.y + a2.y)};
  Vec2 pointA = transformPoint(xfA, f.localPoint);
  Vec2 pointB = transformPoint(xfB, proxyB.p[cache.indexB[0]]);
  float sep = dot(sub(pointB, pointA), normal);
  if (sep < 0.0f) f.axis = neg(f.axis);
  return f;
}

// b2FindMinSeparation, S/distance.c:1018-1083
F2D_HD float findMinSeparation(const SeparationFunction& f, const Proxy& proxyA, const Sweep& sweepA,
                               const Proxy& proxyB, const Sweep& sweepB, int* indexA, int* indexB, float t) {
  Xf xfA = sweepTransform(sweepA, t);
  Xf xfB = sweepTransform(sweepB, t);
  if (f.type == kPointsType) {
    Vec2 axisA = invRotate(xfA.q, f.axis);
    Vec2 axisB = invRotate(xfB.q, neg(f.axis));
    *indexA = findSupport(proxyA, axisA);
    *indexB = findSupport(proxyB, axisB);
    Vec2 pointA = transformPoint(xfA, proxyA.p[*indexA]);
    Vec2 pointB = transformPoint(xfB, proxyB.p[*indexB]);
    return dot(sub(pointB, pointA), f.axis);
  }
  if (f.type == kFaceAType) {
    Vec2 normal = rotate(xfA.q, f.axis);
    Vec2 pointA = transformPoint(xfA, f.localPoint);
    Vec2 axisB = invRotate(xfB.q, neg(normal));
    *indexA = -1;
    *indexB = findSupport(proxyB, axisB);
    Vec2 pointB = transformPoint(xfB, proxyB.p[*indexB]);
    return dot(sub(pointB, pointA), normal);
  }
  Vec2 normal = rotate(xfB.q, f.axis);
  Vec2 pointB = transformPoint(xfB, f.localPoint);
  Vec2 axisA = invRotate(xfA.q, neg(normal));
  *indexB = -1;
  *indexA = findSupport(proxyA, axisA);
  Vec2 pointA = transformPoint(xfA, proxyA.p[*indexA]);
  return dot(sub(pointA, pointB), normal);
}

// b2EvaluateSeparation, S/distance.c:1086-1133
F2D_HD float evaluateSeparation(const SeparationFunction& f, const Proxy& proxyA, const Sweep& sweepA,
                                const Proxy& proxyB, const Sweep& sweepB, int indexA, int indexB, float t) {
  Xf xfA = sweepTransform(sweepA, t);
  Xf xfB = sweepTransform(sweepB, t);
  if (f.type == kPointsType) {
    Vec2 pointA = transformPoint(xfA, proxyA.p[indexA]);
    Vec2 pointB = transformPoint(xfB, proxyB.p[indexB]);
    return dot(sub(pointB, pointA), f.axis);
  }
  if (f.type == kFaceAType) {
    Vec2 normal = rotate(xfA.q, f.axis);
    Vec2 pointA = transformPoint(xfA, f.localPoint);
    Vec2 pointB = transformPoint(xfB, proxyB.p[indexB]);
    return dot(sub(pointB, pointA), normal);
  }
  Vec2 normal = rotate(xfB.q, f.axis);
  Vec2 pointB = transformPoint(xfB, f.localPoint);
  Vec2 pointA = transformPoint(xfA, proxyA.p[indexA]);
  return dot(sub(pointA, pointB), normal);
}

// b2TimeOfImpact, S/distance.c:1137-1414: the fraction of the sweeps in [0, tMax]
// at which the proxies first come within linear slop of touching.
F2D_HD float timeOfImpact(const Proxy& proxyA, const Sweep& sweepA, const Proxy& proxyB, const Sweep& sweepB,
                          float tMax) {
  float fraction = tMax;
  float totalRadius = proxyA.radius + proxyB.radius;
  float target = maxf(kLinearSlop, totalRadius - kLinearSlop);
  float tolerance = 0.25f * kLinearSlop;
  float t1 = 0.0f;
  const int kMaxIterations = 20;
  int distanceIterations = 0;
  SimplexCache cache;
  cache.count = 0;
  for (;;) {
    Xf xfA = sweepTransform(sweepA, t1);
    Xf xfB = sweepTransform(sweepB, t1);
    float distance = shapeDistance(cache, proxyA, proxyB, xfA, xfB);
    distanceIterations += 1;
    if (distance <= 0.0f) {
      fraction = 0.0f;  // overlapped: continuous collision gives up
      break;
    }
    if (distance <= target + tolerance) {
      fraction = t1;
      break;
    }
    SeparationFunction fcn = makeSeparation(cache, proxyA, sweepA, proxyB, sweepB, t1);
    bool done = false;
    float t2 = tMax;
    int pushBackIterations = 0;
    for (;;) {
      int indexA, indexB;
      float s2 = findMinSeparation(fcn, proxyA, sweepA, proxyB, sweepB, &indexA, &indexB, t2);
      if (s2 > target + tolerance) {
        fraction = tMax;  // separated at the end of the sweep
        done = true;
        break;
      }
      if (s2 > target - tolerance) {
        t1 = t2;  // advance the sweeps
        break;
      }
      float s1 = evaluateSeparation(fcn, proxyA, sweepA, proxyB, sweepB, indexA, indexB, t1);
      if (s1 < target - tolerance) {
        fraction = t1;
        done = true;
        break;
      }
      if (s1 <= target + tolerance) {
        fraction = t1;
        done = true;
        break;
      }
      // 1D root of separation(t) = target: alternate bisection and secant steps
      int rootIterationCount = 0;
      float a1 = t1, a2 = t2;
      for (;;) {
        float t;
        if (rootIterationCount & 1) {
          t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
        } else {
          t = 0.5f * (a1 + a2);
        }
        rootIterationCount += 1;
        float s = evaluateSeparation(fcn, proxyA, sweepA, proxyB, sweepB, indexA, indexB, t);
        if (absf(s - target) < tolerance) {
          t2 = t;
          break;
        }
        if (s > target) {
          a1 = t;
          s1 = s;
        } else {
          a2 = t;
          s2 = s;
        }
        if (rootIterationCount == 50) break;
      }
      pushBackIterations += 1;
      if (pushBackIterations == kMaxPolygonVertices) break;
    }
    if (done) break;
    if (distanceIterations == kMaxIterations) {
      fraction = t1;  // root finder stuck
      break;
    }
  }
  return fraction;
}

}  // namespace f2d
