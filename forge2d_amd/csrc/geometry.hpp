// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Host-side shape construction and mass properties with the reference's exact
// arithmetic, so a scene built here is bit-identical to one built by the
// reference: b2MakeBox/b2MakeOffsetRoundedBox S/geometry.c:143-218,
// b2ComputeHull S/hull.c:87-251, b2MakePolygon S/geometry.c:56-82,
// b2ComputePolygonCentroid :27-54, masses :220-399, extents S/shape.c:722-790.
#pragma once

#include "layout.hpp"

namespace f2d {

struct Hull {
  Vec2 points[kMaxPolygonVertices];
  int count;
};

struct MassData {
  float mass;
  Vec2 center;
  float rotationalInertia;
};

inline ShapeGeom makeOffsetRoundedBox(float hx, float hy, Vec2 center, Rot rotation, float radius) {
  Xf xf = {center, rotation};
  ShapeGeom g = {};
  g.type = kPolygon;
  g.count = 4;
  g.v[0] = transformPoint(xf, {-hx, -hy});
  g.v[1] = transformPoint(xf, {hx, -hy});
  g.v[2] = transformPoint(xf, {hx, hy});
  g.v[3] = transformPoint(xf, {-hx, hy});
  g.n[0] = rotate(xf.q, {0.0f, -1.0f});
  g.n[1] = rotate(xf.q, {1.0f, 0.0f});
  g.n[2] = rotate(xf.q, {0.0f, 1.0f});
  g.n[3] = rotate(xf.q, {-1.0f, 0.0f});
  g.radius = radius;
  g.centroid = xf.p;
  return g;
}

inline ShapeGeom makeBox(float hx, float hy) {
  ShapeGeom g = {};
  g.type = kPolygon;
  g.count = 4;
  g.v[0] = {-hx, -hy};
  g.v[1] = {hx, -hy};
  g.v[2] = {hx, hy};
  g.v[3] = {-hx, hy};
  g.n[0] = {0.0f, -1.0f};
  g.n[1] = {1.0f, 0.0f};
  g.n[2] = {0.0f, 1.0f};
  g.n[3] = {-1.0f, 0.0f};
  g.radius = 0.0f;
  g.centroid = {0.0f, 0.0f};
  return g;
}

inline Hull recurseHull(Vec2 p1, Vec2 p2, const Vec2* ps, int count) {
  Hull hull;
  hull.count = 0;
  if (count == 0) return hull;
  Vec2 e = normalize(sub(p2, p1));
  Vec2 rightPoints[kMaxPolygonVertices];
  int rightCount = 0;
  int bestIndex = 0;
  float bestDistance = cross(sub(ps[bestIndex], p1), e);
  if (bestDistance > 0.0f) rightPoints[rightCount++] = ps[bestIndex];
  for (int i = 1; i < count; ++i) {
    float distance = cross(sub(ps[i], p1), e);
    if (distance > bestDistance) {
      bestIndex = i;
      bestDistance = distance;
    }
    if (distance > 0.0f) rightPoints[rightCount++] = ps[i];
  }
  if (bestDistance < 2.0f * kLinearSlop) return hull;
  Vec2 bestPoint = ps[bestIndex];
  Hull hull1 = recurseHull(p1, bestPoint, rightPoints, rightCount);
  Hull hull2 = recurseHull(bestPoint, p2, rightPoints, rightCount);
  for (int i = 0; i < hull1.count; ++i) hull.points[hull.count++] = hull1.points[i];
  hull.points[hull.count++] = bestPoint;
  for (int i = 0; i < hull2.count; ++i) hull.points[hull.count++] = hull2.points[i];
  return hull;
}

// quickhull, S/hull.c:87-251; count == 0 signals a degenerate input
inline Hull computeHull(const Vec2* points, int count) {
  Hull hull;
  hull.count = 0;
  if (count < 3 || count > kMaxPolygonVertices) return hull;
  AABB aabb = {{kFltMax, kFltMax}, {-kFltMax, -kFltMax}};
  Vec2 ps[kMaxPolygonVertices];
  int n = 0;
  const float linearSlop = kLinearSlop;
  const float tolSqr = 16.0f * linearSlop * linearSlop;
  for (int i = 0; i < count; ++i) {
    aabb.lo = vmin(aabb.lo, points[i]);
    aabb.hi = vmax(aabb.hi, points[i]);
    Vec2 vi = points[i];
    bool unique = true;
    for (int j = 0; j < i; ++j) {
      if (distanceSq(vi, points[j]) < tolSqr) {
        unique = false;
        break;
      }
    }
    if (unique) ps[n++] = vi;
  }
  if (n < 3) return hull;
  Vec2 c = {0.5f * (aabb.lo.x + aabb.hi.x), 0.5f * (aabb.lo.y + aabb.hi.y)};
  int f1 = 0;
  float dsq1 = distanceSq(c, ps[f1]);
  for (int i = 1; i < n; ++i) {
    float dsq = distanceSq(c, ps[i]);
    if (dsq > dsq1) {
      f1 = i;
      dsq1 = dsq;
    }
  }
  Vec2 p1 = ps[f1];
  ps[f1] = ps[n - 1];
  n = n - 1;
  int f2 = 0;
  float dsq2 = distanceSq(p1, ps[f2]);
  for (int i = 1; i < n; ++i) {
    float dsq = distanceSq(p1, ps[i]);
    if (dsq > dsq2) {
      f2 = i;
      dsq2 = dsq;
    }
  }
  Vec2 p2 = ps[f2];
  ps[f2] = ps[n - 1];
  n = n - 1;
  Vec2 rightPoints[kMaxPolygonVertices - 2];
  int rightCount = 0;
  Vec2 leftPoints[kMaxPolygonVertices - 2];
  int leftCount = 0;
  Vec2 e = normalize(sub(p2, p1));
  for (int i = 0; i < n; ++i) {
    float d = cross(sub(ps[i], p1), e);
    if (d >= 2.0f * linearSlop) {
      rightPoints[rightCount++] = ps[i];
    } else if (d <= -2.0f * linearSlop) {
      leftPoints[leftCount++] = ps[i];
    }
  }
  Hull hull1 = recurseHull(p1, p2, rightPoints, rightCount);
  Hull hull2 = recurseHull(p2, p1, leftPoints, leftCount);
  if (hull1.count == 0 && hull2.count == 0) return hull;
  hull.points[hull.count++] = p1;
  for (int i = 0; i < hull1.count; ++i) hull.points[hull.count++] = hull1.points[i];
  hull.points[hull.count++] = p2;
  for (int i = 0; i < hull2.count; ++i) hull.points[hull.count++] = hull2.points[i];
  bool searching = true;
  while (searching && hull.count > 2) {
    searching = false;
    for (int i = 0; i < hull.count; ++i) {
      int i1 = i;
      int i2 = (i + 1) % hull.count;
      int i3 = (i + 2) % hull.count;
      Vec2 s1 = hull.points[i1];
      Vec2 s2 = hull.points[i2];
      Vec2 s3 = hull.points[i3];
      Vec2 r = normalize(sub(s3, s1));
      float distance = cross(sub(s2, s1), r);
      if (distance <= 2.0f * linearSlop) {
        for (int j = i2; j < hull.count - 1; ++j) hull.points[j] = hull.points[j + 1];
        hull.count -= 1;
        searching = true;
        break;
      }
    }
  }
  if (hull.count < 3) hull.count = 0;
  return hull;
}

inline Vec2 computePolygonCentroid(const Vec2* vertices, int count) {
  Vec2 center = {0.0f, 0.0f};
  float area = 0.0f;
  Vec2 origin = vertices[0];
  const float inv3 = 1.0f / 3.0f;
  for (int i = 1; i < count - 1; ++i) {
    Vec2 e1 = sub(vertices[i], origin);
    Vec2 e2 = sub(vertices[i + 1], origin);
    float a = 0.5f * cross(e1, e2);
    center = mulAdd(center, a * inv3, add(e1, e2));
    area += a;
  }
  float invArea = 1.0f / area;
  center.x *= invArea;
  center.y *= invArea;
  center = add(origin, center);
  return center;
}

inline ShapeGeom makePolygon(const Hull& hull, float radius) {
  if (hull.count < 3) return makeBox(0.5f, 0.5f);
  ShapeGeom g = {};
  g.type = kPolygon;
  g.count = hull.count;
  g.radius = radius;
  for (int i = 0; i < g.count; ++i) g.v[i] = hull.points[i];
  for (int i = 0; i < g.count; ++i) {
    int i1 = i;
    int i2 = i + 1 < g.count ? i + 1 : 0;
    Vec2 edge = sub(g.v[i2], g.v[i1]);
    g.n[i] = normalize(crossVS(edge, 1.0f));
  }
  g.centroid = computePolygonCentroid(g.v, g.count);
  return g;
}

inline MassData circleMass(Vec2 center, float radius, float density) {
  float rr = radius * radius;
  MassData md;
  md.mass = density * kPi * rr;
  md.center = center;
  md.rotationalInertia = md.mass * (0.5f * rr + dot(center, center));
  return md;
}

inline MassData capsuleMass(Vec2 p1, Vec2 p2, float radius, float density) {
  float rr = radius * radius;
  float len = length(sub(p2, p1));
  float ll = len * len;
  float circleMassV = density * (kPi * radius * radius);
  float boxMass = density * (2.0f * radius * len);
  MassData md;
  md.mass = circleMassV + boxMass;
  md.center.x = 0.5f * (p1.x + p2.x);
  md.center.y = 0.5f * (p1.y + p2.y);
  float lc = 4.0f * radius / (3.0f * kPi);
  float h = 0.5f * len;
  float circleInertia = circleMassV * (0.5f * rr + h * h + 2.0f * h * lc);
  float boxInertia = boxMass * (4.0f * rr + ll) / 12.0f;
  md.rotationalInertia = circleInertia + boxInertia;
  md.rotationalInertia += md.mass * dot(md.center, md.center);
  return md;
}

inline MassData polygonMass(const ShapeGeom& g, float density) {
  if (g.count == 1) return circleMass(g.v[0], g.radius, density);
  if (g.count == 2) return capsuleMass(g.v[0], g.v[1], g.radius, density);
  Vec2 vertices[kMaxPolygonVertices] = {};
  int count = g.count;
  float radius = g.radius;
  if (radius > 0.0f) {
    float sqrt2 = 1.412f;
    for (int i = 0; i < count; ++i) {
      int j = i == 0 ? count - 1 : i - 1;
      Vec2 n1 = g.n[j];
      Vec2 n2 = g.n[i];
      Vec2 mid = normalize(add(n1, n2));
      vertices[i] = mulAdd(g.v[i], sqrt2 * radius, mid);
    }
  } else {
    for (int i = 0; i < count; ++i) vertices[i] = g.v[i];
  }
  Vec2 center = {0.0f, 0.0f};
  float area = 0.0f;
  float rotationalInertia = 0.0f;
  Vec2 r = vertices[0];
  const float inv3 = 1.0f / 3.0f;
  for (int i = 1; i < count - 1; ++i) {
    Vec2 e1 = sub(vertices[i], r);
    Vec2 e2 = sub(vertices[i + 1], r);
    float D = cross(e1, e2);
    float triangleArea = 0.5f * D;
    area += triangleArea;
    center = mulAdd(center, triangleArea * inv3, add(e1, e2));
    float ex1 = e1.x, ey1 = e1.y;
    float ex2 = e2.x, ey2 = e2.y;
    float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
    float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
    rotationalInertia += (0.25f * inv3 * D) * (intx2 + inty2);
  }
  MassData md;
  md.mass = density * area;
  float invArea = 1.0f / area;
  center.x *= invArea;
  center.y *= invArea;
  md.center = add(r, center);
  md.rotationalInertia = density * rotationalInertia;
  md.rotationalInertia += md.mass * (dot(md.center, md.center) - dot(center, center));
  return md;
}

// b2ComputeShapeMass, S/shape.c:706-720
inline MassData shapeMass(const ShapeGeom& g, float density) {
  if (g.type == kCapsule) return capsuleMass(g.v[0], g.v[1], g.radius, density);
  if (g.type == kCircle) return circleMass(g.v[0], g.radius, density);
  if (g.type == kPolygon) return polygonMass(g, density);
  return {0.0f, {0.0f, 0.0f}, 0.0f};
}

// b2ComputeShapeExtent, S/shape.c:722-790
inline void shapeExtent(const ShapeGeom& g, Vec2 localCenter, float* minExtent, float* maxExtent) {
  *minExtent = 0.0f;
  *maxExtent = 0.0f;
  if (g.type == kCapsule) {
    float radius = g.radius;
    *minExtent = radius;
    Vec2 c1 = sub(g.v[0], localCenter);
    Vec2 c2 = sub(g.v[1], localCenter);
    *maxExtent = sqrtf(maxf(lengthSq(c1), lengthSq(c2))) + radius;
  } else if (g.type == kCircle) {
    float radius = g.radius;
    *minExtent = radius;
    *maxExtent = length(sub(g.v[0], localCenter)) + radius;
  } else if (g.type == kPolygon) {
    float mn = kHuge;
    float maxExtentSqr = 0.0f;
    for (int i = 0; i < g.count; ++i) {
      Vec2 v = g.v[i];
      float planeOffset = dot(g.n[i], sub(v, g.centroid));
      mn = minf(mn, planeOffset);
      float distanceSqr = lengthSq(sub(v, localCenter));
      maxExtentSqr = maxf(maxExtentSqr, distanceSqr);
    }
    *minExtent = mn + g.radius;
    *maxExtent = sqrtf(maxExtentSqr) + g.radius;
  } else {
    *minExtent = 0.0f;
    Vec2 c1 = sub(g.v[0], localCenter);
    Vec2 c2 = sub(g.v[1], localCenter);
    *maxExtent = sqrtf(maxf(lengthSq(c1), lengthSq(c2)));
  }
}

}  // namespace f2d
