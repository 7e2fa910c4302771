// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// World queries on a device-resident world: ray casts against each shape type,
// GJK distance with witness points, the conservative-advancement shape cast that
// rounded polygons are ray cast with, and the explosion's projected perimeter —
// restated from S/geometry.c, S/distance.c and S/shape.c in the reference's
// operation order, so that hit fractions, points and normals agree bit for bit.
#pragma once

#include "toi.hpp"

namespace f2d {

// b2RayCastInput, I/collision.h:31-38
struct RayInput {
  Vec2 origin, translation;
  float maxFraction;
};

// b2CastOutput, I/collision.h:75-90 (iterations are not reported)
struct CastOutput {
  Vec2 normal, point;
  float fraction;
  bool hit;
};

// b2DistanceOutput, I/collision.h:478-488 (without the statistics)
struct DistanceOut {
  Vec2 pointA, pointB, normal;
  float distance;
};

// b2InvTransformPoint, I/math_functions.h:509-514
F2D_HD Vec2 invTransformPoint(Xf t, Vec2 p) {
  float vx = p.x - t.p.x;
  float vy = p.y - t.p.y;
  return {t.q.c * vx + t.q.s * vy, -t.q.s * vx + t.q.c * vy};
}

F2D_HD CastOutput noHit() {
  CastOutput o;
  o.normal = {0.0f, 0.0f};
  o.point = {0.0f, 0.0f};
  o.fraction = 0.0f;
  o.hit = false;
  return o;
}

// b2RayCastCircle, S/geometry.c:506-580: closest approach of the ray's line to the
// centre, then Pythagoras; a ray that starts inside reports an initial overlap
// (hit at fraction 0, point = origin)
F2D_HD CastOutput rayCastCircle(const RayInput& in, Vec2 center, float r) {
  CastOutput o = noHit();
  Vec2 s = sub(in.origin, center);
  float rr = r * r;
  float len;
  Vec2 d = lengthAndNormalize(&len, in.translation);
  if (len == 0.0f) {
    if (lengthSq(s) < rr) {
      o.point = in.origin;
      o.hit = true;
    }
    return o;
  }
  float t = -dot(s, d);
  Vec2 c = mulAdd(s, t, d);
  float cc = dot(c, c);
  if (cc > rr) return o;
  float h = sqrtf(rr - cc);
  float fraction = t - h;
  if (fraction < 0.0f || in.maxFraction * len < fraction) {
    if (lengthSq(s) < rr) {
      o.point = in.origin;
      o.hit = true;
    }
    return o;
  }
  Vec2 hitPoint = mulAdd(s, fraction, d);
  o.fraction = fraction / len;
  o.normal = normalize(hitPoint);
  o.point = mulAdd(center, r, o.normal);
  o.hit = true;
  return o;
}

// b2RayCastCapsule, S/geometry.c:582-716: the two side lines by Cramer's rule,
// the end caps as circles
F2D_HD CastOutput rayCastCapsule(const RayInput& in, Vec2 v1, Vec2 v2, float radius) {
  CastOutput o = noHit();
  float capLen;
  Vec2 a = lengthAndNormalize(&capLen, sub(v2, v1));
  if (capLen < kFltEpsilon) return rayCastCircle(in, v1, radius);
  Vec2 p1 = in.origin;
  Vec2 d = in.translation;
  Vec2 q = sub(p1, v1);
  float qa = dot(q, a);
  Vec2 qp = mulAdd(q, -qa, a);
  if (dot(qp, qp) < radius * radius) {
    if (qa < 0.0f) return rayCastCircle(in, v1, radius);
    if (qa > capLen) return rayCastCircle(in, v2, radius);
    o.point = in.origin;  // starts inside the capsule
    o.hit = true;
    return o;
  }
  Vec2 n = {a.y, -a.x};
  float rayLen;
  Vec2 u = lengthAndNormalize(&rayLen, d);
  float den = -a.x * u.y + u.x * a.y;
  if (-kFltEpsilon < den && den < kFltEpsilon) return o;  // parallel, outside
  Vec2 b1 = mulSub(q, radius, n);
  Vec2 b2 = mulAdd(q, radius, n);
  float invDen = 1.0f / den;
  float s21 = (a.x * b1.y - b1.x * a.y) * invDen;
  float s22 = (a.x * b2.y - b2.x * a.y) * invDen;
  float s2;
  Vec2 b;
  if (s21 < s22) {
    s2 = s21;
    b = b1;
  } else {
    s2 = s22;
    b = b2;
    n = neg(n);
  }
  if (s2 < 0.0f || in.maxFraction * rayLen < s2) return o;
  float s1 = (-b.x * u.y + u.x * b.y) * invDen;
  if (s1 < 0.0f) return rayCastCircle(in, v1, radius);
  if (capLen < s1) return rayCastCircle(in, v2, radius);
  o.fraction = s2 / rayLen;
  o.point = add(lerp(v1, v2, s1 / capLen), mulSV(radius, n));
  o.normal = n;
  o.hit = true;
  return o;
}

// b2RayCastSegment, S/geometry.c:718-797 (two-sided)
F2D_HD CastOutput rayCastSegment(const RayInput& in, Vec2 v1, Vec2 v2) {
  CastOutput o = noHit();
  Vec2 p1 = in.origin;
  Vec2 d = in.translation;
  float len;
  Vec2 eUnit = lengthAndNormalize(&len, sub(v2, v1));
  if (len == 0.0f) return o;
  Vec2 normal = rightPerp(eUnit);
  float numerator = dot(normal, sub(v1, p1));
  float denominator = dot(normal, d);
  if (denominator == 0.0f) return o;
  float t = numerator / denominator;
  if (t < 0.0f || in.maxFraction < t) return o;
  Vec2 p = mulAdd(p1, t, d);
  float s = dot(sub(p, v1), eUnit);
  if (s < 0.0f || len < s) return o;
  if (numerator > 0.0f) normal = neg(normal);
  o.fraction = t;
  o.point = p;
  o.normal = normal;
  o.hit = true;
  return o;
}

// b2ShapeDistance, S/distance.c:421-606, with all of its outputs. The early exits
// on overlap return the witness points with distance 0 and leave the cache as it
// was; the regular exit updates the cache and, with useRadii, moves the points
// onto the perimeters.
F2D_HD DistanceOut shapeDistanceOut(SimplexCache& cache, const Proxy& proxyA, const Proxy& proxyB, Xf xfA, Xf xfB,
                                    bool useRadii) {
  DistanceOut out;
  out.pointA = {0.0f, 0.0f};
  out.pointB = {0.0f, 0.0f};
  out.normal = {0.0f, 0.0f};
  out.distance = 0.0f;
  Proxy localB;
  {
    Xf t = invMulTransforms(xfA, xfB);
    localB.count = proxyB.count;
    localB.radius = proxyB.radius;
    for (int i = 0; i < localB.count; ++i) localB.p[i] = transformPoint(t, proxyB.p[i]);
  }
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
  for (int iteration = 0; iteration < 20;) {
    const int saveCount = s.count;
    for (int i = 0; i < saveCount; ++i) {
      saveA[i] = s.v[i].indexA;
      saveB[i] = s.v[i].indexB;
    }
    Vec2 d;
    if (s.count == 1) {
      d = neg(s.v[0].w);
    } else if (s.count == 2) {
      d = solveSimplex2(s);
    } else {
      d = solveSimplex3(s);
    }
    if (s.count == 3 || dot(d, d) < kFltEpsilon * kFltEpsilon) {
      Vec2 la, lb;
      witnessPoints(&la, &lb, s);
      out.pointA = transformPoint(xfA, la);
      out.pointB = transformPoint(xfA, lb);
      return out;
    }
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
  Vec2 normal = rotate(xfA.q, normalize(nonUnitNormal));
  Vec2 la, lb;
  witnessPoints(&la, &lb, s);
  out.normal = normal;
  float dx = lb.x - la.x, dy = lb.y - la.y;
  out.distance = sqrtf(dx * dx + dy * dy);
  out.pointA = transformPoint(xfA, la);
  out.pointB = transformPoint(xfA, lb);
  cache.count = s.count;
  for (int i = 0; i < s.count; ++i) {
    cache.indexA[i] = s.v[i].indexA;
    cache.indexB[i] = s.v[i].indexB;
  }
  if (useRadii && out.distance > 0.1f * kLinearSlop) {
    out.distance = maxf(0.0f, out.distance - proxyA.radius - proxyB.radius);
    out.pointA = mulAdd(out.pointA, proxyA.radius, normal);
    out.pointB = mulSub(out.pointB, proxyB.radius, normal);
  }
  return out;
}

// b2ShapeCast, S/distance.c:608-700: conservative advancement of proxy B along
// translationB towards proxy A (without encroachment)
F2D_HD CastOutput shapeCastPair(const Proxy& proxyA, Xf xfA, const Proxy& proxyB, Xf xfB, Vec2 translationB,
                                float maxFraction) {
  const float linearSlop = kLinearSlop;
  const float totalRadius = proxyA.radius + proxyB.radius;
  const float target = maxf(linearSlop, totalRadius - linearSlop);
  const float tolerance = 0.25f * linearSlop;
  SimplexCache cache;
  cache.count = 0;
  float fraction = 0.0f;
  Xf xfBi = xfB;
  CastOutput o = noHit();
  for (int iteration = 0; iteration < 20; ++iteration) {
    DistanceOut d = shapeDistanceOut(cache, proxyA, proxyB, xfA, xfBi, false);
    if (d.distance < target + tolerance) {
      if (iteration == 0) {
        // initial overlap: a common point between the perimeters
        o.hit = true;
        Vec2 c1 = mulAdd(d.pointA, proxyA.radius, d.normal);
        Vec2 c2 = mulAdd(d.pointB, -proxyB.radius, d.normal);
        o.point = lerp(c1, c2, 0.5f);
        return o;
      }
      o.fraction = fraction;
      o.point = mulAdd(d.pointA, proxyA.radius, d.normal);
      o.normal = d.normal;
      o.hit = true;
      return o;
    }
    float denominator = dot(translationB, d.normal);
    if (denominator >= 0.0f) return o;  // not approaching
    fraction += (target - d.distance) / denominator;
    if (fraction >= maxFraction) return o;
    xfBi.p = mulAdd(xfB.p, fraction, translationB);
  }
  return o;
}

// b2RayCastPolygon, S/geometry.c:799-891: half-space clipping relative to the
// first vertex; a rounded polygon is ray cast as a shape cast of the origin point
F2D_HD CastOutput rayCastPolygon(const RayInput& in, const ShapeGeom& g) {
  const int count = mini(g.count, kMaxPolygonVertices);
  if (g.radius == 0.0f) {
    Vec2 base = g.v[0];
    Vec2 p1 = sub(in.origin, base);
    Vec2 d = in.translation;
    float lower = 0.0f, upper = in.maxFraction;
    int index = -1;
    CastOutput o = noHit();
    for (int i = 0; i < count; ++i) {
      Vec2 vertex = sub(g.v[i], base);
      float numerator = dot(g.n[i], sub(vertex, p1));
      float denominator = dot(g.n[i], d);
      if (denominator == 0.0f) {
        if (numerator < 0.0f) return o;
      } else {
        if (denominator < 0.0f && numerator < lower * denominator) {
          lower = numerator / denominator;  // the ray enters this half-space
          index = i;
        } else if (denominator > 0.0f && numerator < upper * denominator) {
          upper = numerator / denominator;  // the ray leaves this half-space
        }
      }
      if (upper < lower) return o;
    }
    if (index >= 0) {
      o.fraction = lower;
      o.normal = g.n[index];
      o.point = mulAdd(in.origin, lower, d);
    } else {
      o.point = in.origin;  // initial overlap
    }
    o.hit = true;
    return o;
  }
  Proxy a;
  a.count = count;
  for (int i = 0; i < count; ++i) a.p[i] = g.v[i];
  a.radius = g.radius;
  Proxy b;
  b.count = 1;
  b.p[0] = in.origin;
  b.radius = 0.0f;
  const Xf identity = {{0.0f, 0.0f}, {1.0f, 0.0f}};
  return shapeCastPair(a, identity, b, identity, in.translation, in.maxFraction);
}

// b2RayCastShape, S/shape.c:792-820: the ray in the shape's frame, the hit back in
// world space
F2D_HD CastOutput rayCastShape(const RayInput& in, const ShapeGeom& g, Xf xf) {
  RayInput local = in;
  local.origin = invTransformPoint(xf, in.origin);
  local.translation = invRotate(xf.q, in.translation);
  CastOutput o;
  if (g.type == kCapsule) {
    o = rayCastCapsule(local, g.v[0], g.v[1], g.radius);
  } else if (g.type == kCircle) {
    o = rayCastCircle(local, g.v[0], g.radius);
  } else if (g.type == kPolygon) {
    o = rayCastPolygon(local, g);
  } else if (g.type == kSegment) {
    o = rayCastSegment(local, g.v[0], g.v[1]);
  } else {
    return noHit();
  }
  o.point = transformPoint(xf, o.point);
  o.normal = rotate(xf.q, o.normal);
  return o;
}

// b2GetShapeProjectedPerimeter, S/shape.c:656-700: the shape's extent across an
// infinite line through its frame
F2D_HD float projectedPerimeter(const ShapeGeom& g, Vec2 line) {
  if (g.type == kCapsule) {
    Vec2 axis = sub(g.v[1], g.v[0]);
    return absf(dot(axis, line)) + 2.0f * g.radius;
  }
  if (g.type == kCircle) return 2.0f * g.radius;
  if (g.type == kPolygon) {
    const int count = mini(g.count, kMaxPolygonVertices);
    float value = dot(g.v[0], line);
    float lower = value, upper = value;
    for (int i = 1; i < count; ++i) {
      value = dot(g.v[i], line);
      lower = minf(lower, value);
      upper = maxf(upper, value);
    }
    return (upper - lower) + 2.0f * g.radius;
  }
  if (g.type == kSegment) return absf(dot(g.v[1], line) - dot(g.v[0], line));
  return 0.0f;
}

// b2ShouldQueryCollide, S/shape.h:132-135: the shape's category against the query's
// mask and the shape's mask against the query's category (64-bit masks as two words)
F2D_HD bool shouldQuery(const Shape& s, uint32_t catLo, uint32_t catHi, uint32_t maskLo, uint32_t maskHi) {
  return ((s.catLo & maskLo) | (s.catHi & maskHi)) != 0u && ((s.maskLo & catLo) | (s.maskHi & catHi)) != 0u;
}

}  // namespace f2d
