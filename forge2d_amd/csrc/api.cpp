// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// C-ABI of the MI355X world step.
//
// Two entry families, both plain C (scalars, pointers, sizes):
//  * f2d_*  : the flat per-world ABI of the reference's own C shim
//             (F/native/wasm/f2d_shim.c), same names, argument meaning and id
//             packing, for the subset the step path needs;
//  * f2dg_* : the batch ABI: many independent worlds packed into one device
//             batch and stepped by one launch per step.
// Worlds are built on the host exactly as the reference builds them
// (S/body.c:177-330, S/shape.c:36-210, S/joint.c:146-280), committed to the
// device once, stepped there, and mirrored back lazily for getters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "manifold.hpp"
#include "joint.hpp"
#include "query.hpp"
#include "solver.hpp"
#include "world_store.hpp"

extern "C" int f2d_launch_step(const f2d::BatchView* b, const f2d::StepParams* sp, const int* worldList, int nWorlds,
                               hipStream_t stream, const f2d::LaunchConfig* cfg);
extern "C" int f2d_launch_island_command(const f2d::BatchView* b, int world, int op, int arg, hipStream_t stream);
extern "C" int f2d_launch_queries(const f2d::BatchView* b, int* group, const f2d::Query* q, int n,
                                  f2d::QueryHit* hits, int* counts, int cap, hipStream_t stream);
extern "C" int f2d_launch_enlarge_pending(const f2d::BatchView* b, hipStream_t stream);
extern "C" int f2d_launch_explosions(const f2d::BatchView* b, const f2d::Query* q, int n, hipStream_t stream);

namespace f2d {

#define F2D_HIP(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      std::fprintf(stderr, "forge2d_amd: %s failed: %s\n", #expr, hipGetErrorString(e_));    \
      return -(int)e_ - 1000;                                                                \
    }                                                                                        \
  } while (0)

// ---------------------------------------------------------------------------- host construction

void HostWorld::ensureColorWords() {
  int words = ((int)sims.size() + 31) / 32;
  if (words < 1) words = 1;
  if (words == colorWords && (int)colorBits.size() == kOverflowIndex * words) return;
  std::vector<uint32_t> nb((size_t)kOverflowIndex * words, 0u);
  for (int c = 0; c < kOverflowIndex; ++c)
    for (int k = 0; k < colorWords && k < words; ++k) nb[(size_t)c * words + k] = colorBits[(size_t)c * colorWords + k];
  colorBits.swap(nb);
  colorWords = words;
}

// b2CreateBody, S/body.c:177-330
int HostWorld::createBody(int type, Vec2 p, Rot q, Vec2 v, float w, float linDamp, float angDamp, float gravityScale,
                          float sleepThreshold, int flags, bool isAwake) {
  BodySim s = {};
  s.xf.p = p;
  s.xf.q = q;
  s.center = p;
  s.rotation0 = q;
  s.center0 = p;
  s.minExtent = kHuge;
  s.maxExtent = 0.0f;
  s.linearDamping = linDamp;
  s.angularDamping = angDamp;
  s.gravityScale = gravityScale;
  s.type = type;
  s.flags = flags;
  s.headShape = -1;
  s.headContactKey = -1;
  s.contactCount = 0;
  s.headJointKey = -1;
  s.jointCount = 0;
  s.sleepThreshold = sleepThreshold;
  s.sleepTime = 0.0f;
  // solver set and island (b2CreateBody, S/body.c:176-324): static bodies join the
  // static set; a body created awake (or that may not sleep) joins the awake set, a
  // body created asleep a new sleeping set of its own (its velocities are not kept:
  // a sleeping set holds no body states); each non-static body gets an island
  const bool awake = isAwake || (flags & BF_ENABLE_SLEEP) == 0;
  BodyState st = {};
  st.v = v;
  st.w = w;
  st.dq = {1.0f, 0.0f};
  if (type == kStatic || !awake) {
    st.v = {0.0f, 0.0f};
    st.w = 0.0f;
  }
  s.islandId = s.islandPrev = s.islandNext = -1;
  int id = (int)sims.size();
  int sleepSet = -1;
  if (type == kStatic) {
    s.setIndex = kStaticSet;
    s.localIndex = staticBodies++;
  } else if (awake) {
    s.setIndex = kAwakeSet;
    s.localIndex = hdr.awakeBodyCount;
  } else {
    setFree.resize(std::max<size_t>(setFree.size(), (size_t)hdr.setCount + 4));
    sleepSet = poolAlloc(setFree.data(), hdr.setFreeCount, hdr.setCount);
    s.setIndex = sleepSet;
    s.localIndex = 0;
  }
  sims.push_back(s);
  if (type != kStatic && !awake) {
    // b2CreateIslandForBody with the new set (S/body.c:84-97, S/island.c:18-55): the
    // island is the set's first and only island
    islands.resize(islands.size() + 1);
    islandFree.resize(islands.size());
    awakeIslands.resize(islands.size());
    setFree.resize(std::max(islands.size() + 4, setFree.size()));
    hdr.islandCap = (int)islands.size();
    const int isl = poolAlloc(islandFree.data(), hdr.islandFreeCount, hdr.islandCount);
    Island& is = islands[isl];
    is = Island{};
    is.id = isl;
    is.setIndex = sleepSet;
    is.localIndex = 0;
    is.headBody = is.tailBody = id;
    is.bodyCount = 1;
    is.headContact = is.tailContact = -1;
    is.headJoint = is.tailJoint = -1;
    is.parent = -1;
    sims[id].islandId = isl;
  } else if (type != kStatic) {
    awakeBodies.resize(hdr.awakeBodyCount);  // the mirror may hold a committed batch's full-length vector
    awakeBodies.push_back(id);
    hdr.awakeBodyCount = (int)awakeBodies.size();
    // b2CreateIslandForBody, S/body.c:84-97
    islands.resize(islands.size() + 1);
    islandFree.resize(islands.size());
    awakeIslands.resize(islands.size());
    setFree.resize(islands.size() + 4);
    hdr.islandCap = (int)islands.size();
    IslandCtx c = islandCtx();
    int isl = createIsland(c);
    Island& is = islands[isl];
    is.headBody = is.tailBody = id;
    is.bodyCount = 1;
    sims[id].islandId = isl;
  }
  states.push_back(st);
  bodyMass.push_back(0.0f);
  bodyInertia.push_back(0.0f);
  bodyGen.push_back(1);
  bodyFixedRotation.push_back((flags & BF_FIXED_ROTATION) ? 1 : 0);
  ensureColorWords();
  return (int)sims.size() - 1;
}

// b2UpdateBodyMassData, S/body.c:528-631
void HostWorld::updateBodyMass(int bi) {
  BodySim& sim = sims[bi];
  bodyMass[bi] = 0.0f;
  bodyInertia[bi] = 0.0f;
  sim.invMass = 0.0f;
  sim.invI = 0.0f;
  sim.localCenter = {0.0f, 0.0f};
  sim.minExtent = kHuge;
  sim.maxExtent = 0.0f;
  if (sim.type != kDynamic) {
    sim.center = sim.xf.p;
    if (sim.type == kKinematic) {
      for (int s = sim.headShape; s >= 0; s = shapes[s].next) {
        float mn, mx;
        shapeExtent(geoms[s], {0.0f, 0.0f}, &mn, &mx);
        sim.minExtent = minf(sim.minExtent, mn);
        sim.maxExtent = maxf(sim.maxExtent, mx);
      }
    }
    return;
  }
  Vec2 localCenter = {0.0f, 0.0f};
  for (int s = sim.headShape; s >= 0; s = shapes[s].next) {
    if (shapeDensity[s] == 0.0f) continue;
    MassData md = shapeMass(geoms[s], shapeDensity[s]);
    bodyMass[bi] += md.mass;
    localCenter = mulAdd(localCenter, md.mass, md.center);
    bodyInertia[bi] += md.rotationalInertia;
  }
  if (bodyMass[bi] > 0.0f) {
    sim.invMass = 1.0f / bodyMass[bi];
    localCenter = mulSV(sim.invMass, localCenter);
  }
  if (bodyInertia[bi] > 0.0f && bodyFixedRotation[bi] == 0) {
    bodyInertia[bi] -= bodyMass[bi] * dot(localCenter, localCenter);
    sim.invI = 1.0f / bodyInertia[bi];
  } else {
    bodyInertia[bi] = 0.0f;
    sim.invI = 0.0f;
  }
  Vec2 oldCenter = sim.center;
  sim.localCenter = localCenter;
  sim.center = transformPoint(sim.xf, sim.localCenter);
  sim.center0 = sim.center;
  BodyState& st = states[bi];
  Vec2 deltaLinear = crossSV(st.w, sub(sim.center, oldCenter));
  st.v = add(st.v, deltaLinear);
  for (int s = sim.headShape; s >= 0; s = shapes[s].next) {
    float mn, mx;
    shapeExtent(geoms[s], localCenter, &mn, &mx);
    sim.minExtent = minf(sim.minExtent, mn);
    sim.maxExtent = maxf(sim.maxExtent, mx);
  }
}

// b2CreateShapeInternal + b2CreateShapeProxy, S/shape.c:59-161, 901-910
int HostWorld::createShape(int body, const ShapeGeom& g, float density, float friction, float restitution,
                           float rolling, float tangentSpeed, uint64_t cat, uint64_t mask, int group, int flags,
                           bool invokeContact, bool updateMass) {
  Shape sh = {};
  sh.body = body;
  sh.friction = friction;
  sh.restitution = restitution;
  sh.rollingResistance = rolling;
  sh.tangentSpeed = tangentSpeed;
  sh.catLo = (uint32_t)cat;
  sh.catHi = (uint32_t)(cat >> 32);
  sh.maskLo = (uint32_t)mask;
  sh.maskHi = (uint32_t)(mask >> 32);
  sh.group = group;
  sh.flags = flags & (SF_CONTACT_EVENTS | SF_HIT_EVENTS | SF_PRESOLVE | SF_SENSOR);
  BodySim& sim = sims[body];
  int type = sim.type;
  // b2UpdateShapeAABBs, S/shape.c:36-57
  AABB aabb = computeShapeAABB(g, sim.xf);
  aabb.lo.x -= kSpeculativeDistance;
  aabb.lo.y -= kSpeculativeDistance;
  aabb.hi.x += kSpeculativeDistance;
  aabb.hi.y += kSpeculativeDistance;
  sh.aabb = aabb;
  float margin = type == kStatic ? kSpeculativeDistance : kAabbMargin;
  sh.fat.lo.x = aabb.lo.x - margin;
  sh.fat.lo.y = aabb.lo.y - margin;
  sh.fat.hi.x = aabb.hi.x + margin;
  sh.fat.hi.y = aabb.hi.y + margin;
  // b2BroadPhase_CreateProxy (S/broad_phase.c:89-100): a leaf with the fat box in the
  // body type's tree, inserted with rotations (S/dynamic_tree.c:772-795)
  const int proxyId = treeCreateProxy(treeRef(type, 2), sh.fat, (int)shapes.size());
  sh.proxyKey = (proxyId << 2) | type;
  if (type != kStatic || invokeContact || (flags & SF_SENSOR)) {
    sh.flags |= SF_MOVED;
    sh.moveRank = (1 << 30) + createdSeq++;
  }
  sh.next = sim.headShape;
  int id = (int)shapes.size();
  shapes.push_back(sh);
  geoms.push_back(g);
  shapeDensity.push_back(density);
  sims[body].headShape = id;
  if (updateMass) updateBodyMass(body);
  return id;
}

TreeRef HostWorld::treeRef(int type, int extra) {
  TreeHeader& h = tree[type];
  const size_t need = (size_t)std::max(h.fresh + extra, 1);
  if (tnodes[type].size() < need) {
    const size_t cap = std::max(need, tnodes[type].size() + tnodes[type].size() / 2 + 16);
    tnodes[type].resize(cap);
    tfreed[type].resize(cap);
  }
  h.cap = (int)tnodes[type].size();
  return TreeRef{&h, tnodes[type].data(), tfreed[type].data()};
}

// The step enlarged these leaves (b2BroadPhase_EnlargeProxy, S/broad_phase.c:125-135)
// and left them to the next tree pass; the resulting tree does not depend on the
// order they are applied in.
void HostWorld::applyPendingEnlarges() {
  for (Shape& sh : shapes) {
    if ((sh.flags & SF_TREE_PENDING) == 0) continue;
    sh.flags &= ~SF_TREE_PENDING;
    if (sh.proxyKey < 0) continue;
    treeEnlargeProxy(treeRef(sh.proxyKey & 3), sh.proxyKey >> 2, sh.fat);
  }
}

void HostWorld::computeTreeRanks() {
  std::vector<int> rank(shapes.size(), 0), walk;
  for (int t = 0; t < 3; ++t) {
    walk.resize((size_t)tree[t].fresh + 2);
    if (tree[t].root >= 0 && !tnodes[t].empty()) treeLeafRanks(tnodes[t].data(), tree[t].root, walk.data(), rank.data());
  }
  treeRanks = rank;
}

// b2CreateJoint + b2CreateRevoluteJoint + b2AssignJointColor, S/joint.c:146-280, 507-560,
// S/constraint_graph.c:214-290
int HostWorld::createRevolute(int bodyA, int bodyB, Vec2 la, Vec2 lb, float refAngle, float targetAngle,
                              bool enableSpring, float hertz, float damping, bool enableLimit, float lower, float upper,
                              bool enableMotor, float maxMotorTorque, float motorSpeed, bool collideConnected) {
  Joint j = {};
  // b2AllocId: the last freed id first, else a fresh one (S/id_pool.c)
  int id = (int)joints.size();
  if (!jointFree.empty()) {
    id = jointFree.back();
    jointFree.pop_back();
  }
  j.bodyA = bodyA;
  j.bodyB = bodyB;
  j.collideConnected = collideConnected ? 1 : 0;
  j.prevA = -1;
  j.nextA = sims[bodyA].headJointKey;
  int keyA = (id << 1) | 0;
  if (sims[bodyA].headJointKey >= 0) {
    Joint& hj = joints[sims[bodyA].headJointKey >> 1];
    if ((sims[bodyA].headJointKey & 1) == 0) hj.prevA = keyA; else hj.prevB = keyA;
  }
  sims[bodyA].headJointKey = keyA;
  sims[bodyA].jointCount += 1;
  j.prevB = -1;
  j.nextB = sims[bodyB].headJointKey;
  int keyB = (id << 1) | 1;
  if (sims[bodyB].headJointKey >= 0) {
    Joint& hj = joints[sims[bodyB].headJointKey >> 1];
    if ((sims[bodyB].headJointKey & 1) == 0) hj.prevA = keyB; else hj.prevB = keyB;
  }
  sims[bodyB].headJointKey = keyB;
  sims[bodyB].jointCount += 1;
  j.localOriginAnchorA = la;
  j.localOriginAnchorB = lb;
  j.referenceAngle = clampf(refAngle, -kPi, kPi);
  j.targetAngle = clampf(targetAngle, -kPi, kPi);
  j.hertz = hertz;
  j.dampingRatio = damping;
  j.lowerAngle = lower;
  j.upperAngle = upper;
  j.maxMotorTorque = maxMotorTorque;
  j.motorSpeed = motorSpeed;
  j.enableSpring = enableSpring;
  j.enableLimit = enableLimit;
  j.enableMotor = enableMotor;
  j.type = kRevoluteJoint;
  j.constraintHertz = 60.0f;        // B2_JOINT_CONSTRAINT_HERTZ
  j.constraintDampingRatio = 2.0f;  // B2_JOINT_CONSTRAINT_DAMPING_RATIO
  j.constraintSoftness = {0.0f, 1.0f, 0.0f};
  j.color = -1;
  j.localIndex = -1;
  bool staticA = sims[bodyA].type == kStatic, staticB = sims[bodyB].type == kStatic;
  if (!(staticA && staticB)) {
    ensureColorWords();
    int color = kOverflowIndex;
    auto bit = [&](int c, int b) -> bool { return (colorBits[(size_t)c * colorWords + (b >> 5)] >> (b & 31)) & 1u; };
    auto set = [&](int c, int b) { colorBits[(size_t)c * colorWords + (b >> 5)] |= 1u << (b & 31); };
    if (!staticA && !staticB) {
      for (int i = 0; i < kOverflowIndex; ++i) {
        if (bit(i, bodyA) || bit(i, bodyB)) continue;
        set(i, bodyA);
        set(i, bodyB);
        color = i;
        break;
      }
    } else if (!staticA) {
      for (int i = 0; i < kOverflowIndex; ++i) {
        if (bit(i, bodyA)) continue;
        set(i, bodyA);
        color = i;
        break;
      }
    } else {
      for (int i = 0; i < kOverflowIndex; ++i) {
        if (bit(i, bodyB)) continue;
        set(i, bodyB);
        color = i;
        break;
      }
    }
    j.color = color;
    // appended to the colour's joint array (b2AddJointToGraph, S/constraint_graph.c:278-296)
    int n = 0;
    for (const Joint& o : joints)
      if (o.color == color) ++n;
    j.localIndex = n;
  }
  // joints between two static bodies stay in the static set; the others join the
  // awake set and the island graph, and islands merge at once (b2CreateJoint,
  // S/joint.c:204-309; b2LinkJoint with mergeIslands, S/island.c:289-385)
  j.islandId = j.islandPrev = j.islandNext = -1;
  j.setIndex = (staticA && staticB) ? kStaticSet : kAwakeSet;
  if (id == (int)joints.size()) joints.push_back(j);
  else joints[id] = j;
  if ((int)jointGen.size() < (int)joints.size()) jointGen.resize(joints.size(), 0);
  jointGen[id] = (uint16_t)(jointGen[id] + 1);
  if (j.setIndex == kAwakeSet) {
    IslandCtx c = islandCtx();
    linkJointIslands(c, id);
    mergeAwakeIslands(c);
  }
  return id;
}

// b2CreatePrismaticJoint, S/joint.c:559-610: the same body lists, colour and islands
// as a revolute joint (they do not depend on the joint's parameters), then the
// prismatic parameters; the axis is normalized and the reference angle is not
// clamped.
int HostWorld::createPrismatic(int bodyA, int bodyB, Vec2 la, Vec2 lb, Vec2 axis, float refAngle,
                               float targetTranslation, bool enableSpring, float hertz, float damping,
                               bool enableLimit, float lower, float upper, bool enableMotor, float maxMotorForce,
                               float motorSpeed, bool collideConnected) {
  const int id = createRevolute(bodyA, bodyB, la, lb, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kPrismaticJoint;
  j.localAxisA = normalize(axis);
  j.referenceAngle = refAngle;
  j.targetAngle = targetTranslation;
  j.hertz = hertz;
  j.dampingRatio = damping;
  j.lowerAngle = lower;
  j.upperAngle = upper;
  j.maxMotorTorque = maxMotorForce;
  j.motorSpeed = motorSpeed;
  j.enableSpring = enableSpring;
  j.enableLimit = enableLimit;
  j.enableMotor = enableMotor;
  return id;
}

// b2CreateWeldJoint, S/joint.c:612-649 (reference angle not clamped)
int HostWorld::createWeld(int bodyA, int bodyB, Vec2 la, Vec2 lb, float refAngle, float linearHertz,
                          float angularHertz, float linearDamping, float angularDamping, bool collideConnected) {
  const int id = createRevolute(bodyA, bodyB, la, lb, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kWeldJoint;
  j.referenceAngle = refAngle;
  j.hertz = linearHertz;
  j.dampingRatio = linearDamping;
  j.angularHertz = angularHertz;
  j.angularDampingRatio = angularDamping;
  return id;
}

// b2CreateWheelJoint, S/joint.c:651-700 (axis normalized)
int HostWorld::createWheel(int bodyA, int bodyB, Vec2 la, Vec2 lb, Vec2 axis, bool enableSpring, float hertz,
                           float damping, bool enableLimit, float lower, float upper, bool enableMotor,
                           float maxMotorTorque, float motorSpeed, bool collideConnected) {
  const int id = createRevolute(bodyA, bodyB, la, lb, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kWheelJoint;
  j.localAxisA = normalize(axis);
  j.lowerAngle = lower;
  j.upperAngle = upper;
  j.maxMotorTorque = maxMotorTorque;
  j.motorSpeed = motorSpeed;
  j.hertz = hertz;
  j.dampingRatio = damping;
  j.enableSpring = enableSpring;
  j.enableLimit = enableLimit;
  j.enableMotor = enableMotor;
  return id;
}

// b2CreateDistanceJoint, S/joint.c:353-405: lengths clamped below by the linear slop
// (B2_LINEAR_SLOP, 0.005 m), the upper bound of the range at least the given minimum
int HostWorld::createDistance(int bodyA, int bodyB, Vec2 la, Vec2 lb, float length, bool enableSpring, float hertz,
                              float damping, bool enableLimit, float minLength, float maxLength, bool enableMotor,
                              float maxMotorForce, float motorSpeed, bool collideConnected) {
  const int id = createRevolute(bodyA, bodyB, la, lb, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kDistanceJoint;
  j.targetAngle = maxf(length, kLinearSlop);
  j.hertz = hertz;
  j.dampingRatio = damping;
  j.lowerAngle = maxf(minLength, kLinearSlop);
  j.upperAngle = maxf(minLength, maxLength);
  j.maxMotorTorque = maxMotorForce;
  j.motorSpeed = motorSpeed;
  j.enableSpring = enableSpring;
  j.enableLimit = enableLimit;
  j.enableMotor = enableMotor;
  return id;
}

// b2CreateMotorJoint, S/joint.c:406-442: anchors at the body origins; maximum force and
// torque as given (the setters clamp them at 0), correction factor clamped to [0, 1]
int HostWorld::createMotor(int bodyA, int bodyB, Vec2 linearOffset, float angularOffset, float maxForce,
                           float maxTorque, float correctionFactor, bool collideConnected) {
  const int id = createRevolute(bodyA, bodyB, {0.0f, 0.0f}, {0.0f, 0.0f}, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, false, 0.0f, 0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kMotorJoint;
  j.localAxisA = linearOffset;
  j.referenceAngle = angularOffset;
  j.maxMotorTorque = maxForce;
  j.angularHertz = maxTorque;
  j.angularDampingRatio = clampf(correctionFactor, 0.0f, 1.0f);
  return id;
}

// b2CreateMouseJoint, S/joint.c:444-478: the target in each body's frame as the local
// anchors (transforms of the moment of creation)
int HostWorld::createMouse(int bodyA, int bodyB, Vec2 target, float hertz, float damping, float maxForce,
                           bool collideConnected) {
  const Vec2 la = invTransformPoint(sims[bodyA].xf, target);
  const Vec2 lb = invTransformPoint(sims[bodyB].xf, target);
  const int id = createRevolute(bodyA, bodyB, la, lb, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, collideConnected);
  Joint& j = joints[id];
  j.type = kMouseJoint;
  j.localAxisA = target;
  j.hertz = hertz;
  j.dampingRatio = damping;
  j.maxMotorTorque = maxForce;
  return id;
}

// b2CreateFilterJoint, S/joint.c:480-505: no constraint, collideConnected always false
int HostWorld::createFilter(int bodyA, int bodyB) {
  const int id = createRevolute(bodyA, bodyB, {0.0f, 0.0f}, {0.0f, 0.0f}, 0.0f, 0.0f, false, 0.0f, 0.0f, false, 0.0f,
                                0.0f, false, 0.0f, 0.0f, false);
  joints[id].type = kFilterJoint;
  return id;
}

// b2DestroyJointInternal, S/joint.c:702-805, for a joint of the awake set: out of both
// bodies' joint lists and its island (b2UnlinkJoint, S/island.c:384-425, which counts
// the removal for the island split), out of its colour (b2RemoveJointFromGraph,
// S/constraint_graph.c:292-322: the colour's last joint takes its index), its id back
// to the pool. Both bodies are awake, so waking them changes nothing. Joints of a
// sleeping or the static set are refused (-1): their set arrays are not restated.
int HostWorld::destroyJoint(int id) {
  Joint& j = joints[id];
  if (j.setIndex != kAwakeSet) return -1;
  auto unlinkEdge = [&](int body, int edge, int prevKey, int nextKey) {
    if (prevKey >= 0) {
      Joint& p = joints[prevKey >> 1];
      if ((prevKey & 1) == 0) p.nextA = nextKey; else p.nextB = nextKey;
    }
    if (nextKey >= 0) {
      Joint& n = joints[nextKey >> 1];
      if ((nextKey & 1) == 0) n.prevA = prevKey; else n.prevB = prevKey;
    }
    if (sims[body].headJointKey == ((id << 1) | edge)) sims[body].headJointKey = nextKey;
    sims[body].jointCount -= 1;
  };
  unlinkEdge(j.bodyA, 0, j.prevA, j.nextA);
  unlinkEdge(j.bodyB, 1, j.prevB, j.nextB);
  if (j.islandId >= 0) {
    Island& is = islands[j.islandId];
    if (j.islandPrev >= 0) joints[j.islandPrev].islandNext = j.islandNext;
    if (j.islandNext >= 0) joints[j.islandNext].islandPrev = j.islandPrev;
    if (is.headJoint == id) is.headJoint = j.islandNext;
    if (is.tailJoint == id) is.tailJoint = j.islandPrev;
    is.jointCount -= 1;
    is.removeCount += 1;
    j.islandId = j.islandPrev = j.islandNext = -1;
  }
  if (j.color >= 0) {
    const int color = j.color;
    if (color != kOverflowIndex) {
      ensureColorWords();
      colorBits[(size_t)color * colorWords + (j.bodyA >> 5)] &= ~(1u << (j.bodyA & 31));
      colorBits[(size_t)color * colorWords + (j.bodyB >> 5)] &= ~(1u << (j.bodyB & 31));
    }
    int last = -1;
    for (const Joint& o : joints) last += o.color == color ? 1 : 0;
    if (j.localIndex != last) {
      for (Joint& o : joints) {
        if (&o != &j && o.color == color && o.localIndex == last) {
          o.localIndex = j.localIndex;
          break;
        }
      }
    }
  }
  j.setIndex = -1;
  j.localIndex = -1;
  j.color = -1;
  j.prevA = j.nextA = j.prevB = j.nextB = -1;
  jointFree.push_back(id);
  return 0;
}

IslandCtx HostWorld::islandCtx() {
  // the list-walk bounds read the world's current sizes
  hdr.bodyCount = (int)sims.size();
  hdr.jointCount = (int)joints.size();
  hdr.contactCap = std::max(hdr.contactCap, (int)contacts.size());
  IslandCtx c;
  c.h = &hdr;
  c.islands = islands.data();
  c.islandFree = islandFree.data();
  c.awakeIslands = awakeIslands.data();
  c.setFree = setFree.data();
  c.awakeBodies = awakeBodies.data();
  c.moveOrder = moveOrder.data();
  c.awakeBits = nullptr;
  c.stack = nullptr;
  c.seeds = nullptr;
  c.sims = sims.data();
  c.contacts = contacts.data();
  c.links = links.data();
  c.joints = joints.data();
  return c;
}

// ---------------------------------------------------------------------------- scene spec

#pragma pack(push, 1)
struct SpecWorld {
  int32_t body_count, shape_count, joint_count, flags;
  float gx, gy, restitution_threshold, hit_event_threshold, contact_hertz, contact_damping_ratio,
      max_contact_push_speed, maximum_linear_speed;
};
struct SpecBody {
  int32_t type;
  float px, py, qc, qs, vx, vy, w, lin_damp, ang_damp, gravity_scale, sleep_threshold;
  int32_t flags;
};
struct SpecShape {
  int32_t body, kind;
  float friction, restitution, rolling, tangent_speed, density;
  uint32_t cat_lo, cat_hi, mask_lo, mask_hi;
  int32_t group, flags, npts;
  float radius;
  float data[16];
};
struct SpecJoint {
  int32_t body_a, body_b;
  float lax, lay, lbx, lby, ref_angle;
  int32_t enable_limit;
  float lower, upper;
  int32_t enable_motor;
  float motor_speed, max_motor_torque;
  int32_t enable_spring;
  float hertz, damping;
  int32_t collide_connected;
  float draw_size;
};
#pragma pack(pop)
static_assert(sizeof(SpecWorld) == 48 && sizeof(SpecBody) == 52, "spec layout");
static_assert(sizeof(SpecShape) == 124 && sizeof(SpecJoint) == 72, "spec layout");

static void initWorldHeader(HostWorld& hw, const SpecWorld& sw) {
  WorldHeader& h = hw.hdr;
  h.flags = sw.flags;
  h.gravity = {sw.gx, sw.gy};
  h.restitutionThreshold = sw.restitution_threshold;
  h.hitEventThreshold = sw.hit_event_threshold;
  h.maxLinearSpeed = sw.maximum_linear_speed;
  h.contactHertz = sw.contact_hertz;
  h.contactDampingRatio = sw.contact_damping_ratio;
  h.maxContactPushSpeed = sw.max_contact_push_speed;
  h.jointHertz = 60.0f;
  h.jointDampingRatio = 2.0f;
  // b2CreateWorld, S/world.c:138-187: the static, disabled and awake sets exist
  h.setCount = kFirstSleepingSet;
  h.splitIslandId = -1;
}

// Shape kinds of the spec: 0 circle, 1 capsule, 2 segment, 3 polygon (hull of points), 4 box
static int buildSpecShape(HostWorld& hw, const SpecShape& ss, std::string& err) {
  ShapeGeom g = {};
  const float* d = ss.data;
  switch (ss.kind) {
    case 0:
      g.type = kCircle;
      g.v[0] = {d[0], d[1]};
      g.radius = ss.radius;
      break;
    case 1: {
      Vec2 c1 = {d[0], d[1]}, c2 = {d[2], d[3]};
      // b2CreateCapsuleShape, S/shape.c:192-201: a degenerate capsule becomes a circle
      if (distanceSq(c1, c2) <= kLinearSlop * kLinearSlop) {
        g.type = kCircle;
        g.v[0] = lerp(c1, c2, 0.5f);
      } else {
        g.type = kCapsule;
        g.v[0] = c1;
        g.v[1] = c2;
      }
      g.radius = ss.radius;
      break;
    }
    case 2:
      g.type = kSegment;
      g.v[0] = {d[0], d[1]};
      g.v[1] = {d[2], d[3]};
      break;
    case 3: {
      Hull hull = computeHull(reinterpret_cast<const Vec2*>(d), ss.npts);
      if (hull.count == 0) {
        err = "degenerate hull";
        return -1;
      }
      g = makePolygon(hull, ss.radius);
      break;
    }
    case 4:
      g = makeOffsetRoundedBox(d[0], d[1], {d[2], d[3]}, {d[4], d[5]}, ss.radius);
      break;
    default:
      err = "bad shape kind";
      return -1;
  }
  if (g.type == kPolygon && g.count < 3) {
    err = "bad polygon";
    return -1;
  }
  int sflags = 0;
  if (ss.flags & 1) sflags |= SF_CONTACT_EVENTS;
  if (ss.flags & 2) sflags |= SF_HIT_EVENTS;
  if (ss.flags & 8) sflags |= SF_PRESOLVE;
  if (ss.flags & 32) sflags |= SF_SENSOR;
  uint64_t cat = (uint64_t)ss.cat_lo | ((uint64_t)ss.cat_hi << 32);
  uint64_t mask = (uint64_t)ss.mask_lo | ((uint64_t)ss.mask_hi << 32);
  return hw.createShape(ss.body, g, ss.density, ss.friction, ss.restitution, ss.rolling, ss.tangent_speed, cat, mask,
                        ss.group, sflags, (ss.flags & 4) != 0, (ss.flags & 16) != 0);
}

static int parseSpec(Batch& b, const unsigned char* buf, long bytes) {
  if (bytes < 16 || std::memcmp(buf, "F2DSCN01", 8) != 0) {
    b.lastError = "bad scene spec magic";
    return -1;
  }
  int32_t count;
  std::memcpy(&count, buf + 8, 4);
  long off = 16;
  b.worlds.resize(count);
  for (int i = 0; i < count; ++i) {
    if (off + (long)sizeof(SpecWorld) > bytes) {
      b.lastError = "truncated scene spec";
      return -1;
    }
    SpecWorld sw;
    std::memcpy(&sw, buf + off, sizeof(sw));
    off += sizeof(sw);
    long need = (long)sw.body_count * sizeof(SpecBody) + (long)sw.shape_count * sizeof(SpecShape) +
                (long)sw.joint_count * sizeof(SpecJoint);
    if (off + need > bytes) {
      b.lastError = "truncated scene spec";
      return -1;
    }
    HostWorld& hw = b.worlds[i];
    initWorldHeader(hw, sw);
    hw.sims.reserve(sw.body_count);
    for (int k = 0; k < sw.body_count; ++k) {
      SpecBody sb;
      std::memcpy(&sb, buf + off, sizeof(sb));
      off += sizeof(sb);
      int flags = 0;
      if (sb.flags & 1) flags |= BF_BULLET;
      if (sb.flags & 2) flags |= BF_FIXED_ROTATION;
      if (sb.flags & 4) flags |= BF_ALLOW_FAST_ROTATION;
      if (sb.flags & 8) flags |= BF_ENABLE_SLEEP;
      hw.createBody(sb.type, {sb.px, sb.py}, {sb.qc, sb.qs}, {sb.vx, sb.vy}, sb.w, sb.lin_damp, sb.ang_damp,
                    sb.gravity_scale, sb.sleep_threshold, flags, (sb.flags & 16) != 0);
    }
    for (int k = 0; k < sw.shape_count; ++k) {
      SpecShape ss;
      std::memcpy(&ss, buf + off, sizeof(ss));
      off += sizeof(ss);
      if (ss.body < 0 || ss.body >= (int)hw.sims.size()) {
        b.lastError = "shape body out of range";
        return -1;
      }
      if (buildSpecShape(hw, ss, b.lastError) < 0) return -1;
    }
    for (int k = 0; k < sw.joint_count; ++k) {
      SpecJoint sj;
      std::memcpy(&sj, buf + off, sizeof(sj));
      off += sizeof(sj);
      hw.createRevolute(sj.body_a, sj.body_b, {sj.lax, sj.lay}, {sj.lbx, sj.lby}, sj.ref_angle, 0.0f,
                        sj.enable_spring != 0, sj.hertz, sj.damping, sj.enable_limit != 0, sj.lower, sj.upper,
                        sj.enable_motor != 0, sj.max_motor_torque, sj.motor_speed, sj.collide_connected != 0);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------- device commit

static int nextPow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

template <typename T>
static int devAlloc(DeviceBatch& d, T** ptr, size_t count) {
  size_t bytes = sizeof(T) * (count > 0 ? count : 1);
  void* p = nullptr;
  F2D_HIP(hipMalloc(&p, bytes));
  d.allocations.push_back(p);
  d.bytes += bytes;
  *ptr = reinterpret_cast<T*>(p);
  return 0;
}

static void freeDevice(Batch& b) {
  for (void* p : b.dev.allocations) (void)hipFree(p);
  b.dev.allocations.clear();
  b.dev.bytes = 0;
  b.dev.view = BatchView{};
  if (b.dWorldList) (void)hipFree(b.dWorldList);
  b.dWorldList = nullptr;
  b.dWorldListCap = 0;
  if (b.dOrder) (void)hipFree(b.dOrder);
  b.dOrder = nullptr;
  for (SizeClass& c : b.classes) {
    if (c.dList) (void)hipFree(c.dList);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    if (c.done) (void)hipEventDestroy(c.done);
    for (hipEvent_t& e : c.tuneEv) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
  }
  b.classes.clear();
  if (b.forkEvent) (void)hipEventDestroy(b.forkEvent);
  b.forkEvent = nullptr;
  if (b.dQueries) (void)hipFree(b.dQueries);
  if (b.dHits) (void)hipFree(b.dHits);
  if (b.dCounts) (void)hipFree(b.dCounts);
  if (b.dGroup) (void)hipFree(b.dGroup);
  b.dGroup = nullptr;
  b.dGroupCap = 0;
  b.dQueries = nullptr;
  b.dHits = nullptr;
  b.dCounts = nullptr;
  b.dQueryCap = b.dHitCap = 0;
  b.committed = false;
}

// Dynamic LDS bytes a step launch may ask for on the current device: the device's
// LDS per workgroup (160 KiB on gfx950, 64 KiB when the query fails) less 4 KiB
// for the step kernels' own static LDS (at most 1.3 KiB), since a dispatch whose
// static plus dynamic LDS exceeds the device limit is refused.
static int deviceLdsLimit() {
  static int limit = 0;
  if (limit == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v >= 65536) {
      limit = v - 4096;
    } else {
      limit = 65536 - 4096;
    }
  }
  return limit;
}

// Lay out all worlds in flat arrays and allocate the device batch.
// Dynamic LDS of a solve launch: the body states and per-step body terms (64 B per
// body) of the class's largest world, its joints, plus room for `lanesPerBody`
// constraint lanes per body (164 B each); a world whose joints or constraints do
// not fit keeps them in HBM.
static int solveLdsBytes(int maxBodies, int maxJoints, float lanesPerBody, int limit) {
  const int bodies = std::max(maxBodies, 1) * (int)(sizeof(BodyState) + sizeof(BodyConst));
  if (bodies > limit) return 0;
  long total = bodies + (long)maxJoints * (long)sizeof(Joint) +
               (long)(lanesPerBody * (float)std::max(maxBodies, 1)) * 4L * CN_FIELD_COUNT;
  total = std::min<long>(total, limit);
  return (int)((total + 255) & ~255L) <= limit ? (int)((total + 255) & ~255L) : limit;
}

// LDS for the tree kernel's work arrays of a world's larger moving tree (worlds that
// need more than `limit` work from HBM)
static int treeLdsBytes(const HostWorld& hw, int limit) {
  long need = 0;
  for (int t = kKinematic; t <= kDynamic; ++t)
    if (hw.tree[t].proxyCount > 0) need = std::max(need, 4 * treeWorkInts(hw.tree[t].fresh, hw.tree[t].proxyCount));
  return need <= limit ? (int)((need + 255) & ~255L) : 0;
}

// lanes of the tree launch: about two proxies per lane per pass
static int treeBlockFor(int proxies) {
  return proxies <= 96 ? 64 : (proxies <= 192 ? 128 : (proxies <= 4096 ? 256 : 1024));
}

static int commit(Batch& b) {
  freeDevice(b);
  long nb = 0, ns = 0, nc = 0, nj = 0, ncol = 0, nscratch = 0, nisl = 0, ntree = 0, nclist = 0;
  for (HostWorld& hw : b.worlds) {
    hw.ensureColorWords();
    // tree regions: a step's rebuild allocates exactly the nodes it frees, so the
    // node count a tree reaches on the host bounds it on the device
    for (int t = 0; t < 3; ++t) {
      hw.treeRef(t, 16);
      hw.tree[t].offset = (int)ntree;
      ntree += hw.tree[t].cap;
    }
    WorldHeader& h = hw.hdr;
    // island region: one island per non-static body at most
    h.islandCap = std::max(1, (int)hw.islands.size());
    hw.islands.resize(h.islandCap);
    hw.islandFree.resize(h.islandCap);
    hw.awakeIslands.resize(h.islandCap);
    hw.setFree.resize(h.islandCap + 4);
    hw.awakeBodies.resize(hw.sims.size());
    hw.moveOrder.resize(hw.sims.size());
    h.islandOffset = (int)nisl;
    nisl += islandRegion(h.islandCap, (int)hw.sims.size()).total;
    h.colorListOffset = (int)nclist;
    nclist += colorListInts((int)hw.sims.size());
    h.bodyOffset = (int)nb;
    h.bodyCount = (int)hw.sims.size();
    h.shapeOffset = (int)ns;
    h.shapeCount = (int)hw.shapes.size();
    int cap = (int)std::ceil(b.contactCapFactor * h.shapeCount) + 64;
    cap = std::max(cap, (int)hw.contacts.size());
    cap = std::max(cap, hw.contactCount + 64);
    h.contactCap = cap;
    h.contactOffset = (int)nc;
    h.freeOffset = (int)nc;
    h.pairOffset = (int)nc;
    h.pairCap = cap;
    h.overflowOffset = (int)nc;
    h.constraintOffset = (int)nc;
    h.jointOffset = (int)nj;
    h.jointCount = (int)hw.joints.size();
    h.colorOffset = (int)ncol;
    h.colorWords = hw.colorWords;
    h.scratchOffset = (int)nscratch;
    int np = nextPow2(std::max(h.shapeCount, 1));
    int pp = nextPow2(cap);
    long sc = std::max<long>({4L * np + 16, 3L * pp + 16, (long)cap + 16, (long)h.jointCount + 80});
    sc = (sc + 3) & ~3L;  // 16-byte aligned world regions (64-bit sort keys live here)
    h.scratchCap = (int)sc;
    h.contactCount = hw.contactCount;
    h.freeCount = (int)hw.freeIds.size();
    h.overflowCount = (int)hw.overflow.size();
    h.stepIndex = hw.stepIndex;
    h.error = 0;
    nb += h.bodyCount;
    ns += h.shapeCount;
    nc += cap;
    nj += h.jointCount;
    ncol += (long)kOverflowIndex * hw.colorWords;
    nscratch += sc;
    hw.contacts.resize(cap);
    for (int k = hw.contactCount; k < cap; ++k) {
      hw.contacts[k] = Contact{};
      hw.contacts[k].shapeA = -1;
      hw.contacts[k].shapeB = -1;
      hw.contacts[k].color = -1;
    }
    hw.links.resize(cap, make_int4(-1, -1, -1, -1));
  }
  if (nc > (1L << 31) / CN_FIELD_COUNT) {
    b.lastError = "batch too large for 32-bit constraint indexing";
    return -1;
  }
  DeviceBatch& d = b.dev;
  BatchView& v = d.view;
  int W = (int)b.worlds.size();
  if (devAlloc(d, &v.worlds, W) || devAlloc(d, &v.sims, nb) || devAlloc(d, &v.states, nb) ||
      devAlloc(d, &v.shapes, ns) || devAlloc(d, &v.geoms, ns) || devAlloc(d, &v.sapBoxes, ns) ||
      devAlloc(d, &v.contacts, nc) || devAlloc(d, &v.freeIds, nc) || devAlloc(d, &v.overflow, nc) ||
      devAlloc(d, &v.pairs, nc) || devAlloc(d, &v.pairShapes, nc) || devAlloc(d, &v.joints, nj) ||
      devAlloc(d, &v.colorBits, ncol) || devAlloc(d, &v.cons, (size_t)nc * CN_FIELD_COUNT) ||
      devAlloc(d, &v.scratch, nscratch) || devAlloc(d, &v.ccolor, nc) || devAlloc(d, &v.staticShapes, ns) || devAlloc(d, &v.sweepOrder, ns) ||
      devAlloc(d, &v.events, 3 * nc) || devAlloc(d, &v.islandData, std::max(nisl, 4L)) ||
      devAlloc(d, &v.contactLinks, nc) || devAlloc(d, &v.treeHdrs, 3L * W) ||
      devAlloc(d, &v.colorLists, std::max(nclist, 1L)) ||
      devAlloc(d, &v.treeNodes, std::max(ntree, 1L)) || devAlloc(d, &v.treeFreed, std::max(ntree, 1L)) ||
      devAlloc(d, &v.treeWork, treeWorkOffset((int)ntree, W, 0) + 1) || devAlloc(d, &v.treeRanks, ns)) {
    b.lastError = "device allocation failed";
    freeDevice(b);
    return -1;
  }
  v.consStride = nc;
  v.worldCount = W;
  b.nBodies = nb;
  b.nShapes = ns;
  b.nContacts = nc;
  b.nJoints = nj;
  b.nColorWords = ncol;
  b.nIsland = std::max(nisl, 4L);
  b.nTreeNodes = std::max(ntree, 1L);
  b.nColorListInts = std::max(nclist, 1L);
  // dynamic LDS per kernel: sort-and-sweep (28 B per shape, power-of-two padded)
  // and new-pair sort (12 B per pair), body states (32 B per body); worlds that do
  // not fit fall back to their global scratch inside the kernel
  int maxShapes = 1, maxBodies = 1;
  for (HostWorld& hw : b.worlds) {
    maxShapes = std::max(maxShapes, (int)hw.shapes.size());
    maxBodies = std::max(maxBodies, (int)hw.sims.size());
  }
  const int kLdsLimit = 65536;
  int np = nextPow2(maxShapes);
  // sort-and-sweep arrays (28 B per shape), the pair-test cache (32 B per shape) and
  // the existing-pair hash set (16 B per shape), as far as they fit
  auto collideBytes = [&](int p) {
    p = std::max(p, 4);
    return 76 * p <= kLdsLimit ? 76 * p : (44 * p <= kLdsLimit ? 44 * p : (28 * p <= kLdsLimit ? 28 * p : 0));
  };
  int collide = collideBytes(np);
  collide = std::max(collide, 12 * 256);
  b.launch.collideLds = collide;
  // worlds whose sort arrays (28 B per shape) exceed the collide LDS sweep over many workgroups
  b.launch.sweepChunks = 28L * std::max(np, 4) > (long)collide ? (maxShapes + 255) / 256 : 0;
  b.launch.treeLds = 0;
  int maxProxies = 0;
  for (const HostWorld& hw : b.worlds) {
    b.launch.treeLds = std::max(b.launch.treeLds, treeLdsBytes(hw, kLdsLimit));
    maxProxies = std::max(maxProxies, std::max(hw.tree[kDynamic].proxyCount, hw.tree[kKinematic].proxyCount));
  }
  b.launch.treeBlock = treeBlockFor(maxProxies);
  int solve = maxBodies * (int)(sizeof(BodyState) + sizeof(BodyConst));
  b.launch.solveLds = solve <= kLdsLimit ? ((solve + 255) & ~255) : 0;
  if (b.launch.block == 0) b.launch.block = maxShapes <= 640 ? 64 : 256;
  // Full-batch launches dispatch the largest worlds first (longest-processing-time
  // order), so the last wave of workgroups holds the smallest worlds and the
  // launch's tail is short. Worlds are independent: the order changes no result.
  {
    std::vector<int> order(W);
    for (int i = 0; i < W; ++i) order[i] = i;
    auto work = [&](int i) {
      const HostWorld& hw = b.worlds[i];
      return (long)hw.shapes.size() * 4 + (long)hw.sims.size() + (long)hw.joints.size() * 4;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    F2D_HIP(hipMalloc(&b.dOrder, sizeof(int) * std::max(W, 1)));
    F2D_HIP(hipMemcpy(b.dOrder, order.data(), sizeof(int) * W, hipMemcpyHostToDevice));
    // size classes by max(bodies, shapes): 64, 128, ..., 8192, and everything larger
    // (the classes of large worlds stay narrow too: one class's launches size their
    // grids and LDS for its largest world)
    constexpr int kClassLimits = 9;
    const int limits[kClassLimits] = {64, 128, 256, 512, 1024, 2048, 4096, 8192, 1 << 30};
    std::vector<std::vector<int>> members(kClassLimits);
    for (int i : order) {
      const HostWorld& hw = b.worlds[i];
      int size = std::max((int)hw.sims.size(), (int)hw.shapes.size());
      int k = 0;
      while (size > limits[k]) ++k;
      members[k].push_back(i);
    }
    b.classOf.assign(W, 0);
    for (int k = 0; k < kClassLimits; ++k) {
      if (members[k].empty()) continue;
      for (int i : members[k]) b.classOf[i] = (int)b.classes.size();
      SizeClass c;
      c.limit = limits[k];
      c.count = (int)members[k].size();
      for (int i : members[k]) {
        c.maxBodies = std::max(c.maxBodies, (int)b.worlds[i].sims.size());
        c.maxShapes = std::max(c.maxShapes, (int)b.worlds[i].shapes.size());
        c.maxJoints = std::max(c.maxJoints, (int)b.worlds[i].joints.size());
      }
      int cnp = nextPow2(std::max(c.maxShapes, 1));
      int cc = collideBytes(cnp);
      c.launch.collideLds = std::max(cc, 12 * 256);
      c.launch.sweepChunks = 28L * std::max(cnp, 4) > (long)c.launch.collideLds ? (c.maxShapes + 255) / 256 : 0;
      // worlds too large for LDS sorts get 64 KiB for the new-pair sort and the
      // contact state records
      if (c.launch.sweepChunks > 0) c.launch.collideLds = 65536;
      // large worlds spread their non-bullet continuous sweeps over workgroups of
      // 64 fast bodies
      c.launch.continuousChunks = c.maxBodies > 1024 ? (c.maxBodies + 63) / 64 : 0;
      // worlds of many thousand bodies run their solve stages over many workgroups
      c.launch.wideChunks = c.maxBodies > 4096 ? (c.maxBodies + 255) / 256 : 0;
      c.launch.solveLds = solveLdsBytes(c.maxBodies, c.maxJoints, 1.0f, kLdsLimit);
      c.launch.treeLds = 0;
      int maxProxies = 0;
      for (int i : members[k]) {
        c.launch.treeLds = std::max(c.launch.treeLds, treeLdsBytes(b.worlds[i], kLdsLimit));
        maxProxies = std::max(maxProxies, std::max(b.worlds[i].tree[kDynamic].proxyCount,
                                                   b.worlds[i].tree[kKinematic].proxyCount));
      }
      c.launch.treeBlock = treeBlockFor(maxProxies);
      // one workgroup holds a whole world: the largest worlds get the most lanes
      c.launch.block = c.limit <= 128 ? 64 : (c.limit <= 512 ? 128 : (c.limit <= 1024 ? 256 : 1024));
      if (b.forcedBlock) c.launch.block = b.forcedBlock;
      F2D_HIP(hipMalloc(&c.dList, sizeof(int) * c.count));
      F2D_HIP(hipMemcpy(c.dList, members[k].data(), sizeof(int) * c.count, hipMemcpyHostToDevice));
      F2D_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
      F2D_HIP(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
      b.classes.push_back(c);
    }
    F2D_HIP(hipEventCreateWithFlags(&b.forkEvent, hipEventDisableTiming));
  }
  b.committed = true;
  return 0;
}

// Launch one step of every world: each size class on its own stream, forked from
// and joined back into the batch's stream, so classes overlap on the device.
static int launchClasses(Batch& b, const StepParams& sp) {
  if (b.classes.size() == 1) {
    SizeClass& c = b.classes[0];
    return f2d_launch_step(&b.dev.view, &sp, c.dList, c.count, b.stream, &c.launch);
  }
  F2D_HIP(hipEventRecord(b.forkEvent, b.stream));
  for (SizeClass& c : b.classes) {
    F2D_HIP(hipStreamWaitEvent(c.stream, b.forkEvent, 0));
    if (f2d_launch_step(&b.dev.view, &sp, c.dList, c.count, c.stream, &c.launch) != 0) return -1;
    F2D_HIP(hipEventRecord(c.done, c.stream));
  }
  for (SizeClass& c : b.classes) F2D_HIP(hipStreamWaitEvent(b.stream, c.done, 0));
  return 0;
}

// Launch one step of every world, the classes one after another on the batch's
// stream (for per-kernel timing with events on that stream).
static int launchClassesSerial(Batch& b, const StepParams& sp) {
  for (SizeClass& c : b.classes) {
    if (f2d_launch_step(&b.dev.view, &sp, c.dList, c.count, b.stream, &c.launch) != 0) return -1;
  }
  return 0;
}

static int upload(Batch& b) {
  if (!b.committed) {
    if (commit(b) != 0) return -1;
  }
  BatchView& v = b.dev.view;
  std::vector<WorldHeader> headers(b.worlds.size());
  long totalContacts = 0;
  for (const HostWorld& hw : b.worlds) totalContacts = std::max<long>(totalContacts, hw.hdr.contactOffset + hw.hdr.contactCap);
  std::vector<int8_t> ccolor(totalContacts, -1);
  long totalShapes = 0;
  for (const HostWorld& hw : b.worlds) totalShapes = std::max<long>(totalShapes, hw.hdr.shapeOffset + hw.hdr.shapeCount);
  std::vector<int> staticShapes(std::max<long>(totalShapes, 1), 0);
  std::vector<int> sweepOrder(std::max<long>(totalShapes, 1), 0);
  long totalIsland = 4;
  for (const HostWorld& hw : b.worlds)
    totalIsland = std::max<long>(totalIsland, hw.hdr.islandOffset + islandRegion(hw.hdr.islandCap, hw.hdr.bodyCount).total);
  std::vector<int> islandData(totalIsland, 0);
  std::vector<int4> links(std::max<long>(totalContacts, 1), make_int4(-1, -1, -1, -1));
  std::vector<TreeHeader> treeHdrs(3 * b.worlds.size());
  std::vector<TreeNode> treeNodes(b.nTreeNodes);
  std::vector<int> treeFreed(b.nTreeNodes, 0);
  std::vector<int> colorLists(b.nColorListInts, 0);
  b.launch.continuous = 0;
  for (size_t i = 0; i < b.worlds.size(); ++i) {
    HostWorld& hw = b.worlds[i];
    WorldHeader& h = hw.hdr;
    hw.computeTreeRanks();
    if (h.flags & (WF_CONTINUOUS | WF_SLEEP)) b.launch.continuous = 1;  // the last kernel also runs island sleep
    h.fastCount = 0;
    h.staticCount = 0;
    for (int k = 0; k < h.shapeCount; ++k) {
      const Shape& sh = hw.shapes[k];
      if (sh.proxyKey >= 0 && (sh.proxyKey & 3) == kStatic) staticShapes[h.shapeOffset + h.staticCount++] = k;
      sweepOrder[h.shapeOffset + k] = k;  // any permutation; the first pair update sorts it
    }
    // continuous collision visits static candidates in the static tree's query order
    std::sort(staticShapes.begin() + h.shapeOffset, staticShapes.begin() + h.shapeOffset + h.staticCount,
              [&](int x, int y) { return hw.treeRanks[x] < hw.treeRanks[y]; });
    h.contactCount = hw.contactCount;
    h.freeCount = (int)hw.freeIds.size();
    h.overflowCount = (int)hw.overflow.size();
    h.stepIndex = hw.stepIndex;
    h.jointCount = (int)hw.joints.size();
    h.mergePending = 1;  // imported or edited islands may carry parent links
    headers[i] = h;
    hipStream_t s = b.stream;
    F2D_HIP(hipMemcpyAsync(v.sims + h.bodyOffset, hw.sims.data(), sizeof(BodySim) * h.bodyCount, hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.states + h.bodyOffset, hw.states.data(), sizeof(BodyState) * h.bodyCount, hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.shapes + h.shapeOffset, hw.shapes.data(), sizeof(Shape) * h.shapeCount, hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.treeRanks + h.shapeOffset, hw.treeRanks.data(), sizeof(int) * h.shapeCount, hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.geoms + h.shapeOffset, hw.geoms.data(), sizeof(ShapeGeom) * h.shapeCount, hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.contacts + h.contactOffset, hw.contacts.data(), sizeof(Contact) * h.contactCap, hipMemcpyHostToDevice, s));
    for (int k = 0; k < h.contactCap; ++k) {
      const Contact& c = hw.contacts[k];
      ccolor[h.contactOffset + k] = (int8_t)(c.shapeA >= 0 ? c.color : -1);
      if (k < (int)hw.links.size()) links[h.contactOffset + k] = hw.links[k];
    }
    {
      IslandRegion r = islandRegion(h.islandCap, h.bodyCount);
      int* base = islandData.data() + h.islandOffset;
      std::memcpy(base + r.islands, hw.islands.data(), sizeof(Island) * h.islandCap);
      std::memcpy(base + r.islandFree, hw.islandFree.data(), sizeof(int) * h.islandFreeCount);
      std::memcpy(base + r.awakeIslands, hw.awakeIslands.data(), sizeof(int) * h.awakeIslandCount);
      std::memcpy(base + r.setFree, hw.setFree.data(), sizeof(int) * h.setFreeCount);
      std::memcpy(base + r.awakeBodies, hw.awakeBodies.data(), sizeof(int) * h.awakeBodyCount);
      std::memcpy(base + r.moveOrder, hw.moveOrder.data(), sizeof(int) * h.moveBodyCount);
    }
    if (!hw.freeIds.empty())
      F2D_HIP(hipMemcpyAsync(v.freeIds + h.freeOffset, hw.freeIds.data(), sizeof(int) * hw.freeIds.size(), hipMemcpyHostToDevice, s));
    if (!hw.overflow.empty())
      F2D_HIP(hipMemcpyAsync(v.overflow + h.overflowOffset, hw.overflow.data(), sizeof(int) * hw.overflow.size(), hipMemcpyHostToDevice, s));
    if (!hw.joints.empty())
      F2D_HIP(hipMemcpyAsync(v.joints + h.jointOffset, hw.joints.data(), sizeof(Joint) * hw.joints.size(), hipMemcpyHostToDevice, s));
    F2D_HIP(hipMemcpyAsync(v.colorBits + h.colorOffset, hw.colorBits.data(), sizeof(uint32_t) * hw.colorBits.size(), hipMemcpyHostToDevice, s));
    {
      int* cl = colorLists.data() + h.colorListOffset;
      for (int c = 0; c < kOverflowIndex; ++c) {
        cl[c] = (int)hw.colorList[c].size();
        std::copy(hw.colorList[c].begin(), hw.colorList[c].end(), cl + kColorListHeader + (long)c * h.bodyCount);
      }
    }
    for (int t = 0; t < 3; ++t) {
      const TreeHeader& th = hw.tree[t];
      treeHdrs[3 * i + t] = th;
      std::copy(hw.tnodes[t].begin(), hw.tnodes[t].begin() + th.cap, treeNodes.begin() + th.offset);
      std::copy(hw.tfreed[t].begin(), hw.tfreed[t].begin() + th.cap, treeFreed.begin() + th.offset);
    }
  }
  F2D_HIP(hipMemcpyAsync(v.colorLists, colorLists.data(), sizeof(int) * colorLists.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.treeHdrs, treeHdrs.data(), sizeof(TreeHeader) * treeHdrs.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.treeNodes, treeNodes.data(), sizeof(TreeNode) * treeNodes.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.treeFreed, treeFreed.data(), sizeof(int) * treeFreed.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.worlds, headers.data(), sizeof(WorldHeader) * headers.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.ccolor, ccolor.data(), ccolor.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.staticShapes, staticShapes.data(), sizeof(int) * totalShapes, hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.sweepOrder, sweepOrder.data(), sizeof(int) * totalShapes, hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.islandData, islandData.data(), sizeof(int) * islandData.size(), hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipMemcpyAsync(v.contactLinks, links.data(), sizeof(int4) * totalContacts, hipMemcpyHostToDevice, b.stream));
  F2D_HIP(hipStreamSynchronize(b.stream));
  for (SizeClass& c : b.classes) c.launch.continuous = b.launch.continuous;
  b.hostDirty = false;
  return 0;
}

// Copy the device state of every world back to the host mirror: one copy per flat
// array, then a scatter into the worlds' vectors.
static int download(Batch& b) {
  if (!b.committed || !b.deviceNewer) return 0;
  BatchView& v = b.dev.view;
  hipStream_t s = b.stream;
  std::vector<WorldHeader> headers(b.worlds.size());
  std::vector<BodySim> sims(b.nBodies);
  std::vector<BodyState> states(b.nBodies);
  std::vector<Shape> shapes(b.nShapes);
  std::vector<Contact> contacts(b.nContacts);
  std::vector<int> freeIds(b.nContacts), overflow(b.nContacts), events(3 * b.nContacts), islandData(b.nIsland);
  std::vector<Joint> joints(b.nJoints);
  std::vector<uint32_t> colorBits(b.nColorWords);
  std::vector<int4> links(b.nContacts);
  std::vector<TreeHeader> treeHdrs(3 * b.worlds.size());
  std::vector<TreeNode> treeNodes(b.nTreeNodes);
  std::vector<int> treeFreed(b.nTreeNodes);
  std::vector<int> colorLists(b.nColorListInts);
  F2D_HIP(hipMemcpyAsync(colorLists.data(), v.colorLists, sizeof(int) * colorLists.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(treeHdrs.data(), v.treeHdrs, sizeof(TreeHeader) * treeHdrs.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(treeNodes.data(), v.treeNodes, sizeof(TreeNode) * treeNodes.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(treeFreed.data(), v.treeFreed, sizeof(int) * treeFreed.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(headers.data(), v.worlds, sizeof(WorldHeader) * headers.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(sims.data(), v.sims, sizeof(BodySim) * sims.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(states.data(), v.states, sizeof(BodyState) * states.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(shapes.data(), v.shapes, sizeof(Shape) * shapes.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(contacts.data(), v.contacts, sizeof(Contact) * contacts.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(freeIds.data(), v.freeIds, sizeof(int) * freeIds.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(overflow.data(), v.overflow, sizeof(int) * overflow.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(events.data(), v.events, sizeof(int) * events.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(islandData.data(), v.islandData, sizeof(int) * islandData.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(joints.data(), v.joints, sizeof(Joint) * joints.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(colorBits.data(), v.colorBits, sizeof(uint32_t) * colorBits.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipMemcpyAsync(links.data(), v.contactLinks, sizeof(int4) * links.size(), hipMemcpyDeviceToHost, s));
  F2D_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < b.worlds.size(); ++i) {
    HostWorld& hw = b.worlds[i];
    const WorldHeader& h = headers[i];
    std::copy(sims.begin() + h.bodyOffset, sims.begin() + h.bodyOffset + h.bodyCount, hw.sims.begin());
    std::copy(states.begin() + h.bodyOffset, states.begin() + h.bodyOffset + h.bodyCount, hw.states.begin());
    std::copy(shapes.begin() + h.shapeOffset, shapes.begin() + h.shapeOffset + h.shapeCount, hw.shapes.begin());
    hw.contacts.assign(contacts.begin() + h.contactOffset, contacts.begin() + h.contactOffset + h.contactCap);
    hw.links.assign(links.begin() + h.contactOffset, links.begin() + h.contactOffset + h.contactCap);
    hw.freeIds.assign(freeIds.begin() + h.freeOffset, freeIds.begin() + h.freeOffset + h.freeCount);
    hw.overflow.assign(overflow.begin() + h.overflowOffset, overflow.begin() + h.overflowOffset + h.overflowCount);
    std::copy(joints.begin() + h.jointOffset, joints.begin() + h.jointOffset + hw.joints.size(), hw.joints.begin());
    std::copy(colorBits.begin() + h.colorOffset, colorBits.begin() + h.colorOffset + hw.colorBits.size(),
              hw.colorBits.begin());
    const int* ev = events.data() + 3L * h.contactOffset;
    hw.eventsRetired = false;
    hw.beginIds.assign(ev, ev + h.beginCount);
    hw.endPairs.assign(ev + h.contactCap, ev + h.contactCap + 2 * h.endCount);
    {
      IslandRegion r = islandRegion(h.islandCap, h.bodyCount);
      const int* base = islandData.data() + h.islandOffset;
      std::memcpy(hw.islands.data(), base + r.islands, sizeof(Island) * h.islandCap);
      std::memcpy(hw.islandFree.data(), base + r.islandFree, sizeof(int) * h.islandCap);
      std::memcpy(hw.awakeIslands.data(), base + r.awakeIslands, sizeof(int) * h.islandCap);
      std::memcpy(hw.setFree.data(), base + r.setFree, sizeof(int) * (h.islandCap + 4));
      std::memcpy(hw.awakeBodies.data(), base + r.awakeBodies, sizeof(int) * h.bodyCount);
      std::memcpy(hw.moveOrder.data(), base + r.moveOrder, sizeof(int) * h.bodyCount);
    }
    for (int t = 0; t < 3; ++t) {
      const TreeHeader& th = treeHdrs[3 * i + t];
      if (th.error) headers[i].error |= ERR_TREE;
      hw.tree[t] = th;
      hw.tnodes[t].assign(treeNodes.begin() + th.offset, treeNodes.begin() + th.offset + th.cap);
      hw.tfreed[t].assign(treeFreed.begin() + th.offset, treeFreed.begin() + th.offset + th.cap);
    }
    // the host mirror holds the trees as the reference has them after the step
    hw.applyPendingEnlarges();
    {
      const int* cl = colorLists.data() + h.colorListOffset;
      for (int c = 0; c < kOverflowIndex; ++c) {
        const int* first = cl + kColorListHeader + (long)c * h.bodyCount;
        hw.colorList[c].assign(first, first + cl[c]);
      }
    }
    hw.hdr = h;
    hw.contactCount = h.contactCount;
    hw.stepIndex = h.stepIndex;
    hw.contactSpeed = h.contactSpeed;
    hw.inv_h = h.inv_h;
  }
  b.deviceNewer = false;
  return 0;
}

static int ensureWorldList(Batch& b, int n) {
  if (n <= b.dWorldListCap) return 0;
  if (b.dWorldList) (void)hipFree(b.dWorldList);
  F2D_HIP(hipMalloc(&b.dWorldList, sizeof(int) * n));
  b.dWorldListCap = n;
  return 0;
}

static StepParams makeParams(float dt, int subSteps, int phases) {
  // S/world.c:743-757
  StepParams sp;
  sp.subSteps = subSteps < 1 ? 1 : subSteps;
  sp.dt = dt;
  if (dt > 0.0f) {
    sp.inv_dt = 1.0f / dt;
    sp.h = dt / sp.subSteps;
    sp.inv_h = sp.subSteps * sp.inv_dt;
  } else {
    sp.inv_dt = 0.0f;
    sp.h = 0.0f;
    sp.inv_h = 0.0f;
  }
  sp.phases = phases;
  return sp;
}

static int prepareForStep(Batch& b) {
  if (b.hostDirty) {
    if (b.deviceNewer) {
      b.lastError = "host and device both modified";
      return -1;
    }
    if (upload(b) != 0) return -1;
  }
  return 0;
}

// ---------------------------------------------------------------------------- snapshots

#pragma pack(push, 1)
struct SnapHeader {
  int32_t magic, version, step, body_count, shape_count, contact_cap, joint_count, overflow_count, move_count,
      free_id_count, color_total, reserved0;
  float contact_speed, inv_h, reserved1, reserved2;
};
struct SnapBody {
  int32_t type, set_index, local_index, flags;
  float px, py, qc, qs, cx, cy, c0x, c0y, q0c, q0s, lcx, lcy, vx, vy, w, fx, fy, torque, inv_mass, inv_i, min_ext,
      max_ext, sleep_time, reserved;
};
struct SnapShape {
  float aabb[4];
  float fat[4];
  int32_t enlarged, proxy_key, body, reserved;
};
struct SnapPoint {
  float px, py, ax, ay, bx, by, sep, ni, ti, tni, nv;
  int32_t id, persisted;
};
struct SnapContact {
  int32_t id, shape_a, shape_b, flags, color, local_index, sim_flags, point_count;
  float friction, restitution, rolling, tangent_speed, nx, ny, rolling_impulse, reserved;
  SnapPoint pts[2];
};
struct SnapJoint {
  int32_t id, color, local_index, body_a, body_b, enable_limit, enable_motor, enable_spring;
  float linear_impulse_x, linear_impulse_y, spring_impulse, motor_impulse, lower_impulse, upper_impulse, ref_angle,
      reserved;
};

// b2TreeNode as a snapshot record (S/dynamic_tree.c:19-50): box, category bits,
// children (a leaf's shape id in child1), parent, height, flags
struct SnapTreeNode {
  float box[4];
  uint64_t cat;
  int32_t child1, child2, parent;
  uint16_t height, flags;
};
static_assert(sizeof(SnapTreeNode) == 40, "snapshot tree node");
#pragma pack(pop)
static_assert(sizeof(SnapHeader) == 64 && sizeof(SnapBody) == 112 && sizeof(SnapShape) == 48, "snap layout");
static_assert(sizeof(SnapContact) == 168 && sizeof(SnapJoint) == 64, "snap layout");

template <typename T>
static void put(std::vector<unsigned char>& out, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  out.insert(out.end(), p, p + sizeof(T));
}

// Category bits of every node of a tree: a leaf's are its shape's, an internal
// node's the union of its children's (what the reference's node records hold).
static void treeCategories(const HostWorld& hw, int t, std::vector<uint64_t>& cat) {
  const TreeHeader& th = hw.tree[t];
  cat.assign((size_t)std::max(th.fresh, 1), 0);
  if (th.root < 0) return;
  std::vector<int> order, walk{th.root};
  while (!walk.empty()) {
    const int i = walk.back();
    walk.pop_back();
    order.push_back(i);
    const TreeNode& n = hw.tnodes[t][i];
    if (!(n.flags & TN_LEAF)) {
      walk.push_back(n.child1);
      walk.push_back(n.child2);
    }
  }
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    const TreeNode& n = hw.tnodes[t][*it];
    if (n.flags & TN_LEAF) {
      const Shape& sh = hw.shapes[n.child1];
      cat[*it] = (uint64_t)sh.catLo | ((uint64_t)sh.catHi << 32);
    } else {
      cat[*it] = cat[n.child1] | cat[n.child2];
    }
  }
}

static void exportWorld(const HostWorld& hw, std::vector<unsigned char>& out) {
  SnapHeader h = {};
  h.magic = 0x50414E53;
  h.version = 3;
  h.step = hw.stepIndex;
  h.body_count = (int)hw.sims.size();
  h.shape_count = (int)hw.shapes.size();
  h.contact_cap = hw.contactCount;
  h.joint_count = (int)hw.joints.size();
  h.overflow_count = (int)hw.overflow.size();
  std::vector<int> colorLists[kGraphColorCount];
  for (int c = 0; c < kOverflowIndex; ++c) colorLists[c] = hw.colorList[c];
  colorLists[kOverflowIndex] = hw.overflow;
  int total = 0;
  for (auto& l : colorLists) total += (int)l.size();
  h.color_total = total;
  std::vector<std::pair<int, int>> moves;
  for (size_t s = 0; s < hw.shapes.size(); ++s)
    if (hw.shapes[s].flags & SF_MOVED) moves.push_back({hw.shapes[s].moveRank, hw.shapes[s].proxyKey});
  std::sort(moves.begin(), moves.end());
  h.move_count = (int)moves.size();
  h.free_id_count = (int)hw.freeIds.size();
  h.contact_speed = hw.contactSpeed;
  h.inv_h = hw.inv_h;
  put(out, h);
  for (size_t i = 0; i < hw.sims.size(); ++i) {
    const BodySim& s = hw.sims[i];
    const BodyState& st = hw.states[i];
    SnapBody sb = {};
    sb.type = s.type;
    sb.set_index = s.setIndex;
    sb.local_index = s.localIndex;
    sb.flags = ((s.flags & BF_IS_FAST) ? 1 : 0) | ((s.flags & BF_BULLET) ? 2 : 0) | ((s.flags & BF_SPEED_CAPPED) ? 4 : 0);
    sb.px = s.xf.p.x;
    sb.py = s.xf.p.y;
    sb.qc = s.xf.q.c;
    sb.qs = s.xf.q.s;
    sb.cx = s.center.x;
    sb.cy = s.center.y;
    sb.c0x = s.center0.x;
    sb.c0y = s.center0.y;
    sb.q0c = s.rotation0.c;
    sb.q0s = s.rotation0.s;
    sb.lcx = s.localCenter.x;
    sb.lcy = s.localCenter.y;
    if (s.setIndex == kAwakeSet) {  // only awake bodies keep a velocity state
      sb.vx = st.v.x;
      sb.vy = st.v.y;
      sb.w = st.w;
    }
    sb.fx = s.force.x;
    sb.fy = s.force.y;
    sb.torque = s.torque;
    sb.inv_mass = s.invMass;
    sb.inv_i = s.invI;
    sb.min_ext = s.minExtent;
    sb.max_ext = s.maxExtent;
    sb.sleep_time = s.sleepTime;
    put(out, sb);
  }
  for (const Shape& s : hw.shapes) {
    SnapShape ss = {};
    ss.aabb[0] = s.aabb.lo.x;
    ss.aabb[1] = s.aabb.lo.y;
    ss.aabb[2] = s.aabb.hi.x;
    ss.aabb[3] = s.aabb.hi.y;
    ss.fat[0] = s.fat.lo.x;
    ss.fat[1] = s.fat.lo.y;
    ss.fat[2] = s.fat.hi.x;
    ss.fat[3] = s.fat.hi.y;
    ss.enlarged = 0;
    ss.proxy_key = s.proxyKey;
    ss.body = s.body;
    put(out, ss);
  }
  for (int i = 0; i < hw.contactCount; ++i) {
    const Contact& c = hw.contacts[i];
    SnapContact sc = {};
    sc.id = c.shapeA >= 0 ? i : -1;
    if (c.shapeA >= 0) {
      sc.shape_a = c.shapeA;
      sc.shape_b = c.shapeB;
      sc.flags = c.flags & CF_REF_MASK;
      sc.color = c.color;
      sc.local_index = c.localIndex;
      sc.sim_flags = c.simFlags;
      sc.point_count = c.m.pointCount;
      sc.friction = c.friction;
      sc.restitution = c.restitution;
      sc.rolling = c.rollingResistance;
      sc.tangent_speed = c.tangentSpeed;
      sc.nx = c.m.normal.x;
      sc.ny = c.m.normal.y;
      sc.rolling_impulse = c.m.rollingImpulse;
      for (int k = 0; k < 2; ++k) {
        const ManifoldPoint& mp = c.m.points[k];
        SnapPoint& sp = sc.pts[k];
        sp.px = mp.point.x;
        sp.py = mp.point.y;
        sp.ax = mp.anchorA.x;
        sp.ay = mp.anchorA.y;
        sp.bx = mp.anchorB.x;
        sp.by = mp.anchorB.y;
        sp.sep = mp.separation;
        sp.ni = mp.normalImpulse;
        sp.ti = mp.tangentImpulse;
        sp.tni = mp.totalNormalImpulse;
        sp.nv = mp.normalVelocity;
        sp.id = mp.id & 0xFFFF;
        sp.persisted = (mp.id >> 16) & 1;
      }
    }
    put(out, sc);
  }
  for (size_t i = 0; i < hw.joints.size(); ++i) {
    const Joint& j = hw.joints[i];
    SnapJoint sj = {};
    if (j.setIndex < 0) {  // a destroyed joint's slot (jointId null in the reference)
      sj.id = -1;
      put(out, sj);
      continue;
    }
    sj.id = (int)i;
    sj.color = j.color;
    sj.local_index = j.localIndex;
    sj.body_a = j.bodyA;
    sj.body_b = j.bodyB;
    sj.enable_limit = j.enableLimit;
    sj.enable_motor = j.enableMotor;
    sj.enable_spring = j.enableSpring;
    sj.linear_impulse_x = j.linearImpulse.x;
    sj.linear_impulse_y = j.linearImpulse.y;
    sj.spring_impulse = j.springImpulse;
    sj.motor_impulse = j.motorImpulse;
    sj.lower_impulse = j.lowerImpulse;
    sj.upper_impulse = j.upperImpulse;
    sj.ref_angle = j.referenceAngle;
    // joint kind: 0 revolute, 1 prismatic, 2 weld, 3 wheel, 4 distance, 5 motor, 6 mouse, 7 filter
    sj.reserved = j.type == kPrismaticJoint ? 1.0f
                  : j.type == kWeldJoint    ? 2.0f
                  : j.type == kWheelJoint   ? 3.0f
                  : j.type == kDistanceJoint ? 4.0f
                  : j.type == kMotorJoint   ? 5.0f
                  : j.type == kMouseJoint   ? 6.0f
                  : j.type == kFilterJoint  ? 7.0f
                                             : 0.0f;
    put(out, sj);
  }
  for (int c = 0; c < kGraphColorCount; ++c) put(out, (int32_t)colorLists[c].size());
  for (int c = 0; c < kGraphColorCount; ++c)
    for (int id : colorLists[c]) put(out, (int32_t)id);
  for (auto& m : moves) put(out, (int32_t)m.second);
  for (int id : hw.freeIds) put(out, (int32_t)id);
  // island section (version 2)
  const WorldHeader& wh = hw.hdr;
  int32_t ih[16] = {wh.islandCount, wh.islandFreeCount, wh.awakeIslandCount, wh.splitIslandId, wh.setCount,
                    wh.setFreeCount, wh.awakeBodyCount, wh.islandCount, wh.setCount, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 16; ++k) put(out, ih[k]);
  for (int k = 0; k < wh.islandCount; ++k) put(out, hw.islands[k]);
  for (int k = 0; k < wh.islandFreeCount; ++k) put(out, (int32_t)hw.islandFree[k]);
  for (int k = 0; k < wh.awakeIslandCount; ++k) put(out, (int32_t)hw.awakeIslands[k]);
  for (int k = 0; k < wh.setFreeCount; ++k) put(out, (int32_t)hw.setFree[k]);
  for (int k = 0; k < wh.awakeBodyCount; ++k) put(out, (int32_t)hw.awakeBodies[k]);
  for (const BodySim& b : hw.sims) {
    int32_t r[4] = {b.islandId, b.islandPrev, b.islandNext, b.headContactKey};
    for (int32_t x : r) put(out, x);
  }
  for (int i = 0; i < hw.contactCount; ++i) {
    const Contact& c = hw.contacts[i];
    int32_t r[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    if (c.shapeA >= 0) {
      const int4 l = i < (int)hw.links.size() ? hw.links[i] : make_int4(-1, -1, -1, kAwakeSet);
      r[0] = l.x;
      r[1] = l.y;
      r[2] = l.z;
      r[3] = l.w;
      r[4] = c.prevA;
      r[5] = c.nextA;
      r[6] = c.prevB;
      r[7] = c.nextB;
    }
    for (int32_t x : r) put(out, x);
  }
  for (const Joint& j : hw.joints) {
    int32_t r[4] = {j.islandId, j.islandPrev, j.islandNext, j.setIndex};
    for (int32_t x : r) put(out, x);
  }
  // tree section (version 3): header, node records of ids [0, fresh), freed-id stack
  std::vector<uint64_t> cat;
  for (int t = 0; t < 3; ++t) {
    const TreeHeader& th = hw.tree[t];
    treeCategories(hw, t, cat);
    int32_t hd[8] = {th.root, th.nodeCount, th.proxyCount, th.fresh, th.freedCount, 0, 0, 0};
    for (int32_t x : hd) put(out, x);
    for (int i = 0; i < th.fresh; ++i) {
      const TreeNode& n = hw.tnodes[t][i];
      SnapTreeNode r = {};
      r.child1 = r.child2 = r.parent = -1;
      if (n.flags & TN_ALLOCATED) {
        r.box[0] = n.box.lo.x;
        r.box[1] = n.box.lo.y;
        r.box[2] = n.box.hi.x;
        r.box[3] = n.box.hi.y;
        r.cat = cat[i];
        r.child1 = n.child1;
        r.child2 = n.child2;
        r.parent = n.parent;
        r.height = n.height;
        r.flags = n.flags;
      }
      put(out, r);
    }
    for (int k = 0; k < th.freedCount; ++k) put(out, (int32_t)hw.tfreed[t][k]);
  }
}

// The tree section of a version-3 snapshot; older snapshots keep the trees this
// world was built with, their leaves set to the snapshot's fat boxes.
static int importTrees(HostWorld& hw, const SnapHeader& h, const unsigned char* buf, long& off, long bytes,
                       std::string& err) {
  for (Shape& sh : hw.shapes) sh.flags &= ~SF_TREE_PENDING;
  if (h.version < 3) {
    for (int t = 0; t < 3; ++t) {
      TreeRef r = hw.treeRef(t);
      for (int i = 0; i < r.h->fresh; ++i) {
        TreeNode& n = r.n[i];
        n.flags &= ~TN_ENLARGED;
        if ((n.flags & TN_ALLOCATED) && (n.flags & TN_LEAF)) n.box = hw.shapes[n.child1].fat;
      }
      // internal boxes from their children, deepest first
      std::vector<int> order, walk;
      if (r.h->root >= 0) walk.push_back(r.h->root);
      while (!walk.empty()) {
        const int i = walk.back();
        walk.pop_back();
        order.push_back(i);
        if (!(r.n[i].flags & TN_LEAF)) {
          walk.push_back(r.n[i].child1);
          walk.push_back(r.n[i].child2);
        }
      }
      for (auto it = order.rbegin(); it != order.rend(); ++it)
        if (!(r.n[*it].flags & TN_LEAF)) refitNode(r.n, *it);
    }
    return 0;
  }
  for (int t = 0; t < 3; ++t) {
    int32_t hd[8];
    if (off + 32 > bytes) {
      err = "snapshot tree section truncated";
      return -1;
    }
    std::memcpy(hd, buf + off, 32);
    off += 32;
    if (hd[3] < 0 || hd[4] < 0 || hd[4] > hd[3] || hd[2] != hw.tree[t].proxyCount ||
        off + (long)sizeof(SnapTreeNode) * hd[3] + 4L * hd[4] > bytes) {
      err = "snapshot tree section does not match this world's proxies";
      return -1;
    }
    TreeHeader& th = hw.tree[t];
    th.root = hd[0];
    th.nodeCount = hd[1];
    th.proxyCount = hd[2];
    th.fresh = hd[3];
    th.freedCount = hd[4];
    th.error = 0;
    TreeRef r = hw.treeRef(t, 16);
    for (int i = 0; i < th.fresh; ++i) {
      SnapTreeNode rec;
      std::memcpy(&rec, buf + off, sizeof(rec));
      off += sizeof(rec);
      TreeNode& n = r.n[i];
      n.box = {{rec.box[0], rec.box[1]}, {rec.box[2], rec.box[3]}};
      n.child1 = rec.child1;
      n.child2 = rec.child2;
      n.parent = rec.parent;
      n.height = rec.height;
      n.flags = rec.flags;
    }
    std::memcpy(r.freed, buf + off, 4L * th.freedCount);
    off += 4L * th.freedCount;
  }
  return 0;
}

// The island section of a version-2 snapshot: island records, id pools, the awake
// set's orders, body sets and island links, contact sets, island links and body-list
// edges, joint sets and island links.
static int importIslands(HostWorld& hw, const SnapHeader& h, const unsigned char* buf, long& off, long bytes,
                         std::string& err) {
  auto readInts = [&](int32_t* dst, long n) -> bool {
    if (n < 0 || off + 4 * n > bytes) return false;
    std::memcpy(dst, buf + off, 4 * n);
    off += 4 * n;
    return true;
  };
  int32_t ih[16];
  if (!readInts(ih, 16)) {
    err = "snapshot island section truncated";
    return -1;
  }
  WorldHeader& wh = hw.hdr;
  const int cap = std::max((int)hw.islands.size(), 1);
  if (ih[0] < 0 || ih[0] > cap || ih[1] < 0 || ih[1] > cap || ih[2] < 0 || ih[2] > cap || ih[5] < 0 ||
      ih[5] > cap + 4 || ih[6] < 0 || ih[6] > (int)hw.sims.size()) {
    err = "snapshot island section exceeds this world's island capacity";
    return -1;
  }
  hw.islands.resize(cap);
  hw.islandFree.resize(cap);
  hw.awakeIslands.resize(cap);
  hw.setFree.resize(cap + 4);
  hw.awakeBodies.resize(hw.sims.size());
  hw.moveOrder.resize(hw.sims.size());
  bool ok = readInts(reinterpret_cast<int32_t*>(hw.islands.data()), 16L * ih[0]) &&
            readInts(hw.islandFree.data(), ih[1]) && readInts(hw.awakeIslands.data(), ih[2]) &&
            readInts(hw.setFree.data(), ih[5]) && readInts(hw.awakeBodies.data(), ih[6]);
  std::vector<int32_t> bl(4L * hw.sims.size()), cl(8L * h.contact_cap), jl(4L * hw.joints.size());
  ok = ok && readInts(bl.data(), (long)bl.size()) && readInts(cl.data(), (long)cl.size()) &&
       readInts(jl.data(), (long)jl.size());
  if (!ok) {
    err = "snapshot island section truncated";
    return -1;
  }
  wh.islandCount = ih[0];
  wh.islandFreeCount = ih[1];
  wh.awakeIslandCount = ih[2];
  wh.splitIslandId = ih[3];
  wh.setCount = ih[4];
  wh.setFreeCount = ih[5];
  wh.awakeBodyCount = ih[6];
  wh.moveBodyCount = 0;
  // body sets and indices come from the body records
  const unsigned char* bodies = buf + sizeof(SnapHeader);
  for (size_t i = 0; i < hw.sims.size(); ++i) {
    SnapBody sb;
    std::memcpy(&sb, bodies + i * sizeof(SnapBody), sizeof(sb));
    BodySim& s = hw.sims[i];
    s.setIndex = sb.set_index;
    s.localIndex = sb.local_index;
    s.islandId = bl[4 * i];
    s.islandPrev = bl[4 * i + 1];
    s.islandNext = bl[4 * i + 2];
    s.headContactKey = bl[4 * i + 3];
  }
  for (int i = 0; i < h.contact_cap; ++i) {
    Contact& c = hw.contacts[i];
    if (c.shapeA < 0) continue;
    const int32_t* r = cl.data() + 8L * i;
    hw.links[i] = make_int4(r[0], r[1], r[2], r[3]);
    c.prevA = r[4];
    c.nextA = r[5];
    c.prevB = r[6];
    c.nextB = r[7];
    if (r[3] != kAwakeSet) c.flags |= CF_OFF_AWAKE;
    if (r[3] >= kFirstSleepingSet) {  // index in its sleeping set
      SnapContact sc;
      std::memcpy(&sc, buf + sizeof(SnapHeader) + hw.sims.size() * sizeof(SnapBody) +
                           hw.shapes.size() * sizeof(SnapShape) + (size_t)i * sizeof(SnapContact), sizeof(sc));
      c.localIndex = sc.local_index;
    }
  }
  for (size_t i = 0; i < hw.joints.size(); ++i) {
    Joint& j = hw.joints[i];
    j.islandId = jl[4 * i];
    j.islandPrev = jl[4 * i + 1];
    j.islandNext = jl[4 * i + 2];
    j.setIndex = jl[4 * i + 3];
  }
  return 0;
}

// Overlay a snapshot onto a world built from the same scene spec.
static int importWorld(HostWorld& hw, const unsigned char* buf, long bytes, std::string& err) {
  if (bytes < (long)sizeof(SnapHeader)) {
    err = "snapshot too short";
    return -1;
  }
  SnapHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (h.magic != 0x50414E53) {
    err = "bad snapshot magic";
    return -1;
  }
  if (h.body_count != (int)hw.sims.size() || h.shape_count != (int)hw.shapes.size() ||
      h.joint_count != (int)hw.joints.size()) {
    err = "snapshot does not match the world's bodies/shapes/joints";
    return -1;
  }
  long off = sizeof(h);
  long need = (long)h.body_count * sizeof(SnapBody) + (long)h.shape_count * sizeof(SnapShape) +
              (long)h.contact_cap * sizeof(SnapContact) + (long)h.joint_count * sizeof(SnapJoint) + 4L * 12;
  if (off + need > bytes) {
    err = "snapshot truncated";
    return -1;
  }
  for (int i = 0; i < h.body_count; ++i) {
    SnapBody sb;
    std::memcpy(&sb, buf + off, sizeof(sb));
    off += sizeof(sb);
    BodySim& s = hw.sims[i];
    BodyState& st = hw.states[i];
    s.xf.p = {sb.px, sb.py};
    s.xf.q = {sb.qc, sb.qs};
    s.center = {sb.cx, sb.cy};
    s.center0 = {sb.c0x, sb.c0y};
    s.rotation0 = {sb.q0c, sb.q0s};
    s.localCenter = {sb.lcx, sb.lcy};
    s.force = {sb.fx, sb.fy};
    s.torque = sb.torque;
    s.invMass = sb.inv_mass;
    s.invI = sb.inv_i;
    s.minExtent = sb.min_ext;
    s.maxExtent = sb.max_ext;
    s.sleepTime = sb.sleep_time;
    s.headContactKey = -1;
    s.contactCount = 0;
    st.v = {sb.vx, sb.vy};
    st.w = sb.w;
    st.dp = {0.0f, 0.0f};
    st.dq = {1.0f, 0.0f};
  }
  for (int i = 0; i < h.shape_count; ++i) {
    SnapShape ss;
    std::memcpy(&ss, buf + off, sizeof(ss));
    off += sizeof(ss);
    Shape& s = hw.shapes[i];
    s.aabb = {{ss.aabb[0], ss.aabb[1]}, {ss.aabb[2], ss.aabb[3]}};
    s.fat = {{ss.fat[0], ss.fat[1]}, {ss.fat[2], ss.fat[3]}};
    if (ss.proxy_key != s.proxyKey) {
      err = "snapshot proxy keys differ from this build's proxy-key mirror";
      return -1;
    }
    s.flags &= ~SF_MOVED;
  }
  hw.contacts.assign(h.contact_cap, Contact{});
  hw.contactCount = h.contact_cap;
  std::vector<int> live;
  for (int i = 0; i < h.contact_cap; ++i) {
    SnapContact sc;
    std::memcpy(&sc, buf + off, sizeof(sc));
    off += sizeof(sc);
    Contact& c = hw.contacts[i];
    c = Contact{};
    c.shapeA = -1;
    c.shapeB = -1;
    c.color = -1;
    c.localIndex = -1;
    if (sc.id < 0) continue;
    c.shapeA = sc.shape_a;
    c.shapeB = sc.shape_b;
    c.flags = sc.flags & (CF_TOUCHING | CF_CONTACT_EVENTS);
    c.simFlags = sc.sim_flags & (SIM_TOUCHING | SIM_HIT_EVENT | SIM_PRESOLVE);
    c.color = sc.color;
    c.friction = sc.friction;
    c.restitution = sc.restitution;
    c.rollingResistance = sc.rolling;
    c.tangentSpeed = sc.tangent_speed;
    c.m.normal = {sc.nx, sc.ny};
    c.m.rollingImpulse = sc.rolling_impulse;
    c.m.pointCount = sc.point_count;
    for (int k = 0; k < 2; ++k) {
      const SnapPoint& sp = sc.pts[k];
      ManifoldPoint& mp = c.m.points[k];
      mp.point = {sp.px, sp.py};
      mp.anchorA = {sp.ax, sp.ay};
      mp.anchorB = {sp.bx, sp.by};
      mp.separation = sp.sep;
      mp.normalImpulse = sp.ni;
      mp.tangentImpulse = sp.ti;
      mp.totalNormalImpulse = sp.tni;
      mp.normalVelocity = sp.nv;
      mp.id = (sp.id & 0xFFFF) | (sp.persisted ? 0x10000 : 0);
    }
    live.push_back(i);
  }
  // body contact lists, built in ascending contact id (list order does not affect results)
  for (int id : live) {
    Contact& c = hw.contacts[id];
    int bodyA = hw.shapes[c.shapeA].body, bodyB = hw.shapes[c.shapeB].body;
    c.bodyA = bodyA;
    c.bodyB = bodyB;
    BodySim& ba = hw.sims[bodyA];
    c.prevA = -1;
    c.nextA = ba.headContactKey;
    if (ba.headContactKey >= 0) {
      Contact& hc = hw.contacts[ba.headContactKey >> 1];
      if ((ba.headContactKey & 1) == 0) hc.prevA = (id << 1); else hc.prevB = (id << 1);
    }
    ba.headContactKey = id << 1;
    ba.contactCount++;
    BodySim& bb = hw.sims[bodyB];
    c.prevB = -1;
    c.nextB = bb.headContactKey;
    if (bb.headContactKey >= 0) {
      Contact& hc = hw.contacts[bb.headContactKey >> 1];
      if ((bb.headContactKey & 1) == 0) hc.prevA = (id << 1) | 1; else hc.prevB = (id << 1) | 1;
    }
    bb.headContactKey = (id << 1) | 1;
    bb.contactCount++;
  }
  for (int i = 0; i < h.joint_count; ++i) {
    SnapJoint sj;
    std::memcpy(&sj, buf + off, sizeof(sj));
    off += sizeof(sj);
    Joint& j = hw.joints[i];
    j.color = sj.color;
    j.localIndex = sj.local_index;
    j.linearImpulse = {sj.linear_impulse_x, sj.linear_impulse_y};
    j.springImpulse = sj.spring_impulse;
    j.motorImpulse = sj.motor_impulse;
    j.lowerImpulse = sj.lower_impulse;
    j.upperImpulse = sj.upper_impulse;
  }
  int counts[kGraphColorCount];
  std::memcpy(counts, buf + off, sizeof(counts));
  off += sizeof(counts);
  long idsTotal = 0;
  for (int c = 0; c < kGraphColorCount; ++c) idsTotal += counts[c];
  if (off + 4 * (idsTotal + h.move_count + h.free_id_count) > bytes) {
    err = "snapshot truncated";
    return -1;
  }
  // colour bitsets from contacts and joints, overflow list in array order
  hw.ensureColorWords();
  std::fill(hw.colorBits.begin(), hw.colorBits.end(), 0u);
  auto setBit = [&](int c, int b) { hw.colorBits[(size_t)c * hw.colorWords + (b >> 5)] |= 1u << (b & 31); };
  hw.overflow.clear();
  for (auto& l : hw.colorList) l.clear();
  for (int c = 0; c < kGraphColorCount; ++c) {
    for (int k = 0; k < counts[c]; ++k) {
      int32_t id;
      std::memcpy(&id, buf + off, 4);
      off += 4;
      Contact& ct = hw.contacts[id];
      if (c == kOverflowIndex) {
        ct.localIndex = (int)hw.overflow.size();
        hw.overflow.push_back(id);
      } else {
        ct.localIndex = (int)hw.colorList[c].size();
        hw.colorList[c].push_back(id);
        if (hw.sims[ct.bodyA].type != kStatic) setBit(c, ct.bodyA);
        if (hw.sims[ct.bodyB].type != kStatic) setBit(c, ct.bodyB);
      }
    }
  }
  for (const Joint& j : hw.joints) {
    if (j.color >= 0 && j.color < kOverflowIndex) {
      if (hw.sims[j.bodyA].type != kStatic) setBit(j.color, j.bodyA);
      if (hw.sims[j.bodyB].type != kStatic) setBit(j.color, j.bodyB);
    }
  }
  for (int k = 0; k < h.move_count; ++k) {
    int32_t key;
    std::memcpy(&key, buf + off, 4);
    off += 4;
    for (Shape& s : hw.shapes) {
      if (s.proxyKey == key) {
        s.flags |= SF_MOVED;
        s.moveRank = k;
        break;
      }
    }
  }
  hw.freeIds.resize(h.free_id_count);
  for (int k = 0; k < h.free_id_count; ++k) {
    std::memcpy(&hw.freeIds[k], buf + off, 4);
    off += 4;
  }
  hw.stepIndex = h.step;
  hw.contactSpeed = h.contact_speed;
  hw.inv_h = h.inv_h;
  hw.links.assign(h.contact_cap, make_int4(-1, -1, -1, -1));
  for (int id : live) hw.links[id] = make_int4(-1, -1, -1, kAwakeSet);
  if (h.version >= 2 && importIslands(hw, h, buf, off, bytes, err) != 0) return -1;
  return importTrees(hw, h, buf, off, bytes, err);
}

// ---------------------------------------------------------------------------- handle tables

static std::mutex g_lock;
struct WorldSlot {
  Batch* batch = nullptr;
  uint16_t generation = 0;
};
static std::vector<WorldSlot> g_worlds;  // f2d_* single-world handles

static Batch* batchOf(uint32_t w) {
  uint32_t index1 = w & 0xFFFF, gen = w >> 16;
  if (index1 == 0 || index1 > g_worlds.size()) return nullptr;
  WorldSlot& s = g_worlds[index1 - 1];
  if (s.batch == nullptr || s.generation != gen) return nullptr;
  return s.batch;
}

static HostWorld* bodyWorld(int32_t index1, uint32_t wg, int* bodyId) {
  uint32_t world0 = wg & 0xFFFF, gen = wg >> 16;
  if (world0 >= g_worlds.size()) return nullptr;
  WorldSlot& s = g_worlds[world0];
  if (!s.batch) return nullptr;
  if (s.batch->deviceNewer && download(*s.batch) != 0) return nullptr;
  HostWorld& hw = s.batch->worlds[0];
  int id = index1 - 1;
  if (id < 0 || id >= (int)hw.sims.size() || hw.bodyGen[id] != gen) return nullptr;
  *bodyId = id;
  return &hw;
}

static Batch* bodyBatch(uint32_t wg) {
  uint32_t world0 = wg & 0xFFFF;
  if (world0 >= g_worlds.size()) return nullptr;
  return g_worlds[world0].batch;
}

}  // namespace f2d

using namespace f2d;

// ============================================================================ C-ABI

extern "C" {

// ---- batch ABI --------------------------------------------------------------

void* f2dg_create_from_spec(const void* spec, long bytes) {
  Batch* b = new Batch();
  if (parseSpec(*b, static_cast<const unsigned char*>(spec), bytes) != 0) {
    std::fprintf(stderr, "forge2d_amd: %s\n", b->lastError.c_str());
    delete b;
    return nullptr;
  }
  return b;
}

void f2dg_destroy(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return;
  freeDevice(*b);
  if (b->ownsStream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int f2dg_world_count(void* h) { return h ? (int)static_cast<Batch*>(h)->worlds.size() : -1; }

int f2dg_body_count(void* h, int world) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  return (int)b->worlds[world].sims.size();
}

int f2dg_shape_count(void* h, int world) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  return (int)b->worlds[world].shapes.size();
}

long f2dg_awake_body_count(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  long n = 0;
  for (auto& w : b->worlds)
    for (auto& s : w.sims) n += s.setIndex == kAwakeSet ? 1 : 0;
  return n;
}

int f2dg_set_contact_capacity_factor(void* h, float factor) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || b->committed || factor <= 0.0f) return -1;
  b->contactCapFactor = factor;
  return 0;
}

int f2dg_set_stream(void* h, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  b->stream = static_cast<hipStream_t>(stream);
  return 0;
}

int f2dg_upload(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  return upload(*b);
}

long f2dg_device_bytes(void* h) {
  Batch* b = static_cast<Batch*>(h);
  return b ? (long)b->dev.bytes : -1;
}

// Step every world of the batch `steps` times: one launch per step, one workgroup per world.
// Lane candidates of a size class for online tuning: 64, 128 and 256 lanes per
// world, 256, 512 and 1024 for worlds of more than 512 shapes or bodies.
static int laneCandidate(const SizeClass& c, int k) {
  static const int small[3] = {64, 128, 256}, large[3] = {256, 512, 1024};
  return c.limit > 512 ? large[k] : small[k];
}

// One online tuning round: three steps, step k with every class at its k-th lane
// candidate, the classes one after another on the batch's stream with HIP events
// around each; then every class keeps its fastest candidate. The steps are
// ordinary steps (the lanes never change results).
static int tuneRound(Batch& b, const StepParams& sp) {
  for (SizeClass& c : b.classes) {
    for (hipEvent_t& e : c.tuneEv) {
      if (!e) F2D_HIP(hipEventCreate(&e));
    }
  }
  for (int k = 0; k < 3; ++k) {
    for (SizeClass& c : b.classes) {
      LaunchConfig cfg = c.launch;
      cfg.block = laneCandidate(c, k);
      F2D_HIP(hipEventRecord(c.tuneEv[2 * k], b.stream));
      if (f2d_launch_step(&b.dev.view, &sp, c.dList, c.count, b.stream, &cfg) != 0) return -1;
      F2D_HIP(hipEventRecord(c.tuneEv[2 * k + 1], b.stream));
    }
  }
  F2D_HIP(hipEventSynchronize(b.classes.back().tuneEv[5]));
  for (SizeClass& c : b.classes) {
    float best = 1e30f;
    for (int k = 0; k < 3; ++k) {
      float ms = 0.0f;
      F2D_HIP(hipEventElapsedTime(&ms, c.tuneEv[2 * k], c.tuneEv[2 * k + 1]));
      if (ms < best) {
        best = ms;
        c.launch.block = laneCandidate(c, k);
      }
    }
  }
  return 0;
}

int f2dg_step(void* h, float dt, int subSteps, int steps, int phases) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (prepareForStep(*b) != 0) return -1;
  if (dt == 0.0f) return 0;
  StepParams sp = makeParams(dt, subSteps, phases == 0 ? PH_ALL : phases);
  for (int s = 0; s < steps;) {
    if (b->tunePeriod > 0 && sp.phases == PH_ALL && b->tuneClock >= b->tuneNext && steps - s >= 3) {
      if (tuneRound(*b, sp) != 0) {
        std::fprintf(stderr, "forge2d_amd: step launch failed in a tuning round\n");
        return -1;
      }
      b->tuneNext = b->tuneClock + b->tunePeriod;
      s += 3;
      b->tuneClock += 3;
      continue;
    }
    int rc = launchClasses(*b, sp);
    if (rc != 0) {
      std::fprintf(stderr, "forge2d_amd: step launch failed (%d)\n", rc);
      return -1;
    }
    ++s;
    if (b->tunePeriod > 0) ++b->tuneClock;
  }
  b->deviceNewer = true;
  return 0;
}

// Online lane tuning for f2dg_step: every `period` steps a three-step trial round
// picks each size class's lanes per world again (0 switches it off).
int f2dg_set_online_tuning(void* h, int period) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || period < 0) return -1;
  b->tunePeriod = period;
  b->tuneClock = 0;
  b->tuneNext = 0;
  return 0;
}

// Step a subset of worlds (serving: the worlds that have pending requests).
int f2dg_step_worlds(void* h, const int* worldIds, int n, float dt, int subSteps, int steps) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || n < 0) return -1;
  if (prepareForStep(*b) != 0) return -1;
  if (n == 0 || dt == 0.0f) return 0;
  for (int i = 0; i < n; ++i) {
    if (worldIds[i] < 0 || worldIds[i] >= (int)b->worlds.size()) return -1;
  }
  if (ensureWorldList(*b, n) != 0) return -1;
  StepParams sp = makeParams(dt, subSteps, PH_ALL);
  const int K = (int)b->classes.size();
  if (K == 0 || (int)b->classOf.size() != (int)b->worlds.size()) {
    F2D_HIP(hipMemcpyAsync(b->dWorldList, worldIds, sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
    for (int s = 0; s < steps; ++s) {
      if (f2d_launch_step(&b->dev.view, &sp, b->dWorldList, n, b->stream, &b->launch) != 0) return -1;
    }
    b->deviceNewer = true;
    return 0;
  }
  // the listed worlds grouped by size class, each group with its class's launch
  // configuration on its class's stream (forked from and joined into the batch's)
  std::vector<int> start(K + 1, 0), fill(K, 0);
  for (int i = 0; i < n; ++i) start[b->classOf[worldIds[i]] + 1] += 1;
  for (int k = 0; k < K; ++k) start[k + 1] += start[k];
  std::vector<int> grouped(n);
  for (int i = 0; i < n; ++i) {
    const int k = b->classOf[worldIds[i]];
    grouped[start[k] + fill[k]++] = worldIds[i];
  }
  F2D_HIP(hipMemcpyAsync(b->dWorldList, grouped.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
  for (int s = 0; s < steps; ++s) {
    F2D_HIP(hipEventRecord(b->forkEvent, b->stream));
    for (int k = 0; k < K; ++k) {
      const int cnt = start[k + 1] - start[k];
      if (cnt == 0) continue;
      SizeClass& c = b->classes[k];
      F2D_HIP(hipStreamWaitEvent(c.stream, b->forkEvent, 0));
      if (f2d_launch_step(&b->dev.view, &sp, b->dWorldList + start[k], cnt, c.stream, &c.launch) != 0) return -1;
      F2D_HIP(hipEventRecord(c.done, c.stream));
    }
    for (int k = 0; k < K; ++k)
      if (start[k + 1] > start[k]) F2D_HIP(hipStreamWaitEvent(b->stream, b->classes[k].done, 0));
  }
  b->deviceNewer = true;
  return 0;
}

// The one-world-at-a-time GPU path: each world gets its own launch per step.
int f2dg_step_each(void* h, float dt, int subSteps, int steps) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (prepareForStep(*b) != 0) return -1;
  int W = (int)b->worlds.size();
  if (ensureWorldList(*b, W) != 0) return -1;
  std::vector<int> ids(W);
  for (int i = 0; i < W; ++i) ids[i] = i;
  F2D_HIP(hipMemcpyAsync(b->dWorldList, ids.data(), sizeof(int) * W, hipMemcpyHostToDevice, b->stream));
  StepParams sp = makeParams(dt, subSteps, PH_ALL);
  for (int i = 0; i < W; ++i) {
    for (int s = 0; s < steps; ++s) {
      if (f2d_launch_step(&b->dev.view, &sp, b->dWorldList + i, 1, b->stream, &b->launch) != 0) return -1;
    }
  }
  b->deviceNewer = true;
  return 0;
}

// Time `steps` batch steps with HIP events on the batch's stream; writes the total ms.
int f2dg_step_timed(void* h, float dt, int subSteps, int steps, float* msOut) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (prepareForStep(*b) != 0) return -1;
  hipEvent_t e0, e1;
  F2D_HIP(hipEventCreate(&e0));
  F2D_HIP(hipEventCreate(&e1));
  StepParams sp = makeParams(dt, subSteps, PH_ALL);
  int W = (int)b->worlds.size();
  (void)W;
  F2D_HIP(hipEventRecord(e0, b->stream));
  for (int s = 0; s < steps; ++s) {
    if (launchClasses(*b, sp) != 0) return -1;
  }
  F2D_HIP(hipEventRecord(e1, b->stream));
  F2D_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  F2D_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *msOut = ms;
  b->deviceNewer = true;
  return 0;
}

// Time `steps` batch steps per kernel: HIP events on the batch's stream bracket the
// collide launch, the solve launch and the continuous-collision launch of every
// step; ms[0..2] = summed milliseconds of each (ms[2] = 0 when no world of the
// batch has continuous collision enabled).
int f2dg_kernel_times(void* h, float dt, int subSteps, int steps, float* ms) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || steps <= 0) return -1;
  if (prepareForStep(*b) != 0) return -1;
  hipEvent_t ev[5];
  for (auto& e : ev) F2D_HIP(hipEventCreate(&e));
  // launch order of a step: collide, tree, solve, continuous; reported as
  // ms[0] collide, ms[1] solve, ms[2] continuous, ms[3] tree
  StepParams sp[4] = {makeParams(dt, subSteps, PH_PAIRS | PH_COLLIDE | PH_STATE | PH_NO_COUNT | PH_NO_TREE),
                      makeParams(dt, subSteps, PH_TREE | PH_NO_COUNT),
                      makeParams(dt, subSteps, PH_SOLVE | PH_FINALIZE | PH_NO_CONTINUOUS),
                      makeParams(dt, subSteps, PH_CONTINUOUS)};
  const int slot[4] = {0, 3, 1, 2};
  float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int s = 0; s < steps; ++s) {
    F2D_HIP(hipEventRecord(ev[0], b->stream));
    for (int k = 0; k < 4; ++k) {
      if (launchClassesSerial(*b, sp[k]) != 0) return -1;
      F2D_HIP(hipEventRecord(ev[k + 1], b->stream));
    }
    F2D_HIP(hipEventSynchronize(ev[4]));
    for (int k = 0; k < 4; ++k) {
      float t = 0.0f;
      F2D_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      sum[slot[k]] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 4; ++k) ms[k] = sum[k];
  b->deviceNewer = true;
  return 0;
}

// Step and time the step's phases as separate launches (size classes one after
// another): ms[0] broad-phase pairs, ms[1] narrow phase + contact state loop,
// ms[2] solver + finalize, ms[3] continuous collision + island sleep. The result
// equals a fused step's.
int f2dg_phase_times(void* h, float dt, int subSteps, int steps, float* ms) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || steps <= 0) return -1;
  if (prepareForStep(*b) != 0) return -1;
  hipEvent_t ev[5];
  for (auto& e : ev) F2D_HIP(hipEventCreate(&e));
  StepParams sp[4] = {makeParams(dt, subSteps, PH_PAIRS | PH_NO_COUNT),
                      makeParams(dt, subSteps, PH_COLLIDE | PH_STATE | PH_NO_COUNT),
                      makeParams(dt, subSteps, PH_SOLVE | PH_FINALIZE | PH_NO_CONTINUOUS),
                      makeParams(dt, subSteps, PH_CONTINUOUS)};
  float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int s = 0; s < steps; ++s) {
    F2D_HIP(hipEventRecord(ev[0], b->stream));
    for (int k = 0; k < 4; ++k) {
      if (launchClassesSerial(*b, sp[k]) != 0) return -1;
      F2D_HIP(hipEventRecord(ev[k + 1], b->stream));
    }
    F2D_HIP(hipEventSynchronize(ev[4]));
    for (int k = 0; k < 4; ++k) {
      float t = 0.0f;
      F2D_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      sum[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 4; ++k) ms[k] = sum[k];
  b->deviceNewer = true;
  return 0;
}

int f2dg_synchronize(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  F2D_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

// Copy device state back to the host mirror; returns the OR of the worlds' error bits.
int f2dg_sync(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (download(*b) != 0) return -1;
  int err = 0;
  for (auto& w : b->worlds) err |= w.hdr.error;
  return err;
}

// OR of the worlds' device error bits, reading only the world headers (no mirror update).
int f2dg_error_bits(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (!b->committed) return 0;
  std::vector<WorldHeader> headers(b->worlds.size());
  std::vector<TreeHeader> trees(3 * b->worlds.size());
  F2D_HIP(hipMemcpyAsync(headers.data(), b->dev.view.worlds, sizeof(WorldHeader) * headers.size(),
                         hipMemcpyDeviceToHost, b->stream));
  F2D_HIP(hipMemcpyAsync(trees.data(), b->dev.view.treeHdrs, sizeof(TreeHeader) * trees.size(),
                         hipMemcpyDeviceToHost, b->stream));
  F2D_HIP(hipStreamSynchronize(b->stream));
  int err = 0;
  for (const WorldHeader& w : headers) err |= w.error;
  for (const TreeHeader& t : trees) err |= t.error ? ERR_TREE : 0;
  return err;
}

// Per-body [px, py, qc, qs, vx, vy, w] of one world (after f2dg_sync).
int f2dg_get_bodies(void* h, int world, float* out) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  if (download(*b) != 0) return -1;
  const HostWorld& hw = b->worlds[world];
  for (size_t i = 0; i < hw.sims.size(); ++i) {
    const BodySim& s = hw.sims[i];
    const BodyState& st = hw.states[i];
    float* o = out + 7 * i;
    o[0] = s.xf.p.x;
    o[1] = s.xf.p.y;
    o[2] = s.xf.q.c;
    o[3] = s.xf.q.s;
    o[4] = s.type == kStatic ? 0.0f : st.v.x;
    o[5] = s.type == kStatic ? 0.0f : st.v.y;
    o[6] = s.type == kStatic ? 0.0f : st.w;
  }
  return (int)hw.sims.size();
}

// Serialize one world's state (snapshot layout). Returns the byte count; writes
// only when cap is large enough.
long f2dg_export_world(void* h, int world, void* out, long cap) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  if (download(*b) != 0) return -1;
  std::vector<unsigned char> buf;
  exportWorld(b->worlds[world], buf);
  if (out && cap >= (long)buf.size()) std::memcpy(out, buf.data(), buf.size());
  return (long)buf.size();
}

int f2dg_import_world(void* h, int world, const void* snap, long bytes) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  if (download(*b) != 0) return -1;
  if (importWorld(b->worlds[world], static_cast<const unsigned char*>(snap), bytes, b->lastError) != 0) {
    std::fprintf(stderr, "forge2d_amd: import: %s\n", b->lastError.c_str());
    return -1;
  }
  b->hostDirty = true;
  if (b->committed) {
    // capacities may need to grow; re-lay out the batch
    freeDevice(*b);
  }
  return 0;
}

const char* f2dg_last_error(void* h) {
  Batch* b = static_cast<Batch*>(h);
  return b ? b->lastError.c_str() : "null batch";
}

// ---------------------------------------------------------------------------- events

// Contact events of a world's last step, in the record layouts of the flat shim's
// getters (F/native/wasm/f2d_shim.c:1186-1300). Begin and end events come from the
// device's serial contact state loop in ascending contact id, as the reference
// pushes them (S/world.c:587-690, S/contact.c:354-372); hit events are gathered
// from the stored manifolds like S/solver.c:1758-1817, colour by colour (within a
// colour in contact id order, the reference uses its colour array order).
static int hitEvents(Batch& b, int world, double* out, uint32_t shapeWg);

static void contactEventCounts(Batch& b, int world, int* out) {
  download(b);
  HostWorld& hw = b.worlds[world];
  out[0] = (int)hw.beginIds.size();
  out[1] = (int)hw.endPairs.size() / 2;
  out[2] = 0;
  out[2] = hitEvents(b, world, nullptr, 0);
}

static int beginEvents(Batch& b, int world, double* out, uint32_t shapeWg) {
  download(b);
  HostWorld& hw = b.worlds[world];
  int n = (int)hw.beginIds.size();
  for (int i = 0; out && i < n; ++i) {
    const Contact& c = hw.contacts[hw.beginIds[i]];
    double* r = out + 13 * i;
    r[0] = c.shapeA + 1;
    r[1] = shapeWg;
    r[2] = c.shapeB + 1;
    r[3] = shapeWg;
    r[4] = c.m.normal.x;
    r[5] = c.m.normal.y;
    r[6] = c.m.pointCount;
    for (int p = 0; p < 2; ++p) {
      r[7 + 3 * p] = c.m.points[p].point.x;
      r[8 + 3 * p] = c.m.points[p].point.y;
      r[9 + 3 * p] = c.m.points[p].separation;
    }
  }
  return n;
}

static int endEvents(Batch& b, int world, double* out, uint32_t shapeWg) {
  download(b);
  HostWorld& hw = b.worlds[world];
  int n = (int)hw.endPairs.size() / 2;
  for (int i = 0; out && i < n; ++i) {
    double* r = out + 4 * i;
    r[0] = hw.endPairs[2 * i] + 1;
    r[1] = shapeWg;
    r[2] = hw.endPairs[2 * i + 1] + 1;
    r[3] = shapeWg;
  }
  return n;
}

static int hitEvents(Batch& b, int world, double* out, uint32_t shapeWg) {
  download(b);
  HostWorld& hw = b.worlds[world];
  if (hw.eventsRetired) return 0;
  const float threshold = hw.hdr.hitEventThreshold;
  int n = 0;
  auto visit = [&](int id) {
    const Contact& c = hw.contacts[id];
    if ((c.simFlags & SIM_HIT_EVENT) == 0) return;
    float approach = threshold;
    Vec2 point = {0.0f, 0.0f};
    bool hit = false;
    for (int k = 0; k < c.m.pointCount; ++k) {
      const ManifoldPoint& mp = c.m.points[k];
      float a = -mp.normalVelocity;
      if (a > approach && mp.totalNormalImpulse > 0.0f) {
        approach = a;
        point = mp.point;
        hit = true;
      }
    }
    if (!hit) return;
    if (out) {
      double* r = out + 9 * n;
      r[0] = c.shapeA + 1;
      r[1] = shapeWg;
      r[2] = c.shapeB + 1;
      r[3] = shapeWg;
      r[4] = point.x;
      r[5] = point.y;
      r[6] = c.m.normal.x;
      r[7] = c.m.normal.y;
      r[8] = approach;
    }
    ++n;
  };
  // colour by colour in array order, the overflow colour last (S/solver.c:1768-1816)
  for (int color = 0; color < kOverflowIndex; ++color)
    for (int id : hw.colorList[color]) visit(id);
  for (int id : hw.overflow) visit(id);
  return n;
}

// Body move events: every body the last step finalized, in the awake set's order
// of that finalize (S/solver.c:605-610); fellAsleep marks the bodies whose island
// went to sleep at the end of the step (S/solver_set.c:196-205).
static int bodyEvents(Batch& b, int world, double* out, int worldIndex0) {
  download(b);
  HostWorld& hw = b.worlds[world];
  const int n = std::min(hw.hdr.moveBodyCount, (int)hw.moveOrder.size());
  if (out) {
    for (int k = 0; k < n; ++k) {
      const int i = hw.moveOrder[k];
      const BodySim& s = hw.sims[i];
      double* r = out + 7 * k;
      r[0] = (double)i + 1;
      r[1] = worldIndex0 < 0 ? 0.0 : (double)((uint32_t)worldIndex0 | ((uint32_t)hw.bodyGen[i] << 16));
      r[2] = s.xf.p.x;
      r[3] = s.xf.p.y;
      r[4] = s.xf.q.c;
      r[5] = s.xf.q.s;
      r[6] = (s.flags & BF_FELL_ASLEEP) ? 1.0 : 0.0;
    }
  }
  return n;
}

int f2dg_contact_event_counts(void* h, int world, int32_t* out) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  int c[3];
  contactEventCounts(*b, world, c);
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  return 0;
}
int f2dg_contact_begin_events(void* h, int world, double* out) {
  Batch* b = static_cast<Batch*>(h);
  return (b && world >= 0 && world < (int)b->worlds.size()) ? beginEvents(*b, world, out, 0) : -1;
}
int f2dg_contact_end_events(void* h, int world, double* out) {
  Batch* b = static_cast<Batch*>(h);
  return (b && world >= 0 && world < (int)b->worlds.size()) ? endEvents(*b, world, out, 0) : -1;
}
int f2dg_contact_hit_events(void* h, int world, double* out) {
  Batch* b = static_cast<Batch*>(h);
  return (b && world >= 0 && world < (int)b->worlds.size()) ? hitEvents(*b, world, out, 0) : -1;
}
int f2dg_body_events(void* h, int world, double* out) {
  Batch* b = static_cast<Batch*>(h);
  return (b && world >= 0 && world < (int)b->worlds.size()) ? bodyEvents(*b, world, out, -1) : -1;
}

int f2dg_world_counters(void* h, int world, int* out) {
  // [bodies, shapes, live contacts, touching contacts, overflow, active colours, step index, errors]
  Batch* b = static_cast<Batch*>(h);
  if (!b || world < 0 || world >= (int)b->worlds.size()) return -1;
  if (download(*b) != 0) return -1;
  const HostWorld& hw = b->worlds[world];
  int live = 0, touching = 0;
  for (int i = 0; i < hw.contactCount; ++i) {
    if (hw.contacts[i].shapeA >= 0) {
      live++;
      touching += (hw.contacts[i].flags & CF_TOUCHING) ? 1 : 0;
    }
  }
  out[0] = (int)hw.sims.size();
  out[1] = (int)hw.shapes.size();
  out[2] = live;
  out[3] = touching;
  out[4] = (int)hw.overflow.size();
  out[5] = hw.hdr.activeColors;
  out[6] = hw.stepIndex;
  out[7] = hw.hdr.error;
  return 0;
}

// Workgroup-size autotuner: for each size class, time `steps` steps of that
// class's worlds with each candidate size (64, 128, 256 lanes per world) and keep
// the fastest. msOut (optional, 3 floats) receives the whole-batch ms per
// candidate summed over classes. The trial steps advance every world by the same
// number of steps; results do not depend on the workgroup size.
int f2dg_autotune(void* h, float dt, int subSteps, int steps, float* msOut) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (prepareForStep(*b) != 0) return -1;
  const int cands[5] = {64, 128, 256, 512, 1024};
  StepParams sp = makeParams(dt, subSteps, PH_ALL);
  hipEvent_t e0, e1;
  F2D_HIP(hipEventCreate(&e0));
  F2D_HIP(hipEventCreate(&e1));
  if (msOut) msOut[0] = msOut[1] = msOut[2] = 0.0f;
  for (SizeClass& c : b->classes) {
    float best = 1e30f;
    int bestBlock = c.launch.block;
    // 512 and 1024 lanes only for worlds of more than 256 shapes or bodies
    const int nCands = c.limit > 256 ? 5 : 3;
    for (int k = 0; k < nCands; ++k) {
      LaunchConfig cfg = c.launch;
      cfg.block = cands[k];
      F2D_HIP(hipEventRecord(e0, b->stream));
      for (int s = 0; s < steps; ++s) {
        if (f2d_launch_step(&b->dev.view, &sp, c.dList, c.count, b->stream, &cfg) != 0) return -1;
      }
      F2D_HIP(hipEventRecord(e1, b->stream));
      F2D_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      F2D_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (msOut && k < 3) msOut[k] += ms;
      if (ms < best) {
        best = ms;
        bestBlock = cands[k];
      }
    }
    c.launch.block = bestBlock;
    // classes with fewer candidates run untimed steps so that every class, and so
    // every world, advances by the same 9 x steps trial steps
    for (int s = 0; s < (5 - nCands) * steps; ++s) {
      if (f2d_launch_step(&b->dev.view, &sp, c.dList, c.count, b->stream, &c.launch) != 0) return -1;
    }
    // LDS room for constraint lanes: more keeps more worlds' constraints on chip,
    // less lets more workgroups share a compute unit; every class runs the same
    // number of trial steps, so all worlds advance equally
    const float lanesPerBody[4] = {0.0f, 0.5f, 1.0f, 2.0f};
    best = 1e30f;
    int bestLds = c.launch.solveLds;
    for (int k = 0; k < 4; ++k) {
      LaunchConfig cfg = c.launch;
      cfg.solveLds = solveLdsBytes(c.maxBodies, c.maxJoints, lanesPerBody[k], 65536);
      F2D_HIP(hipEventRecord(e0, b->stream));
      for (int s = 0; s < steps; ++s) {
        if (f2d_launch_step(&b->dev.view, &sp, c.dList, c.count, b->stream, &cfg) != 0) return -1;
      }
      F2D_HIP(hipEventRecord(e1, b->stream));
      F2D_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      F2D_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bestLds = cfg.solveLds;
      }
    }
    c.launch.solveLds = bestLds;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->deviceNewer = true;
  return b->classes.empty() ? b->launch.block : b->classes.back().launch.block;
}

// Force the solve launch's dynamic LDS bytes of one size class (at least its body
// terms; at most the device's LDS per workgroup).
int f2dg_set_class_solve_lds(void* h, int cls, int bytes) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || cls < 0 || cls >= (int)b->classes.size() || bytes < 0 || bytes > deviceLdsLimit()) return -1;
  SizeClass& c = b->classes[cls];
  const int bodies = std::max(c.maxBodies, 1) * (int)(sizeof(BodyState) + sizeof(BodyConst));
  if (bytes < bodies && bytes != 0) return -1;
  c.launch.solveLds = bytes;
  return 0;
}

// Size classes after commit: per class [limit, worlds, max bodies, max shapes,
// workgroup lanes, collide LDS bytes, solve LDS bytes]; returns the class count.
int f2dg_size_classes(void* h, int* out, int cap) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (!b->committed && upload(*b) != 0) return -1;
  int n = (int)b->classes.size();
  for (int k = 0; k < n && k < cap; ++k) {
    const SizeClass& c = b->classes[k];
    int* o = out + 7 * k;
    o[0] = c.limit;
    o[1] = c.count;
    o[2] = c.maxBodies;
    o[3] = c.maxShapes;
    o[4] = c.launch.block;
    o[5] = c.launch.collideLds;
    o[6] = c.launch.solveLds;
  }
  return n;
}

int f2dg_set_block(void* h, int block) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)) return -1;
  b->launch.block = block;
  b->forcedBlock = block;  // also for classes laid out later
  for (SizeClass& c : b->classes) c.launch.block = block;
  return 0;
}

// Force the workgroup size of one size class (classes as f2dg_size_classes lists them).
int f2dg_set_class_block(void* h, int cls, int block) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)) return -1;
  if (cls < 0 || cls >= (int)b->classes.size()) return -1;
  b->classes[cls].launch.block = block;
  return 0;
}

int f2dg_get_block(void* h) {
  Batch* b = static_cast<Batch*>(h);
  return b ? b->launch.block : -1;
}

// Aggregate counts for the algorithmic byte model (after f2dg_sync):
// [awake bodies, shapes, touching contacts, touching manifold points,
//  touching contacts with restitution, their points, joints, overflow contacts]
int f2dg_batch_stats(void* h, long* out) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return -1;
  if (download(*b) != 0) return -1;
  long v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (const HostWorld& hw : b->worlds) {
    for (const BodySim& s : hw.sims) v[0] += s.setIndex == kAwakeSet ? 1 : 0;
    v[1] += (long)hw.shapes.size();
    v[8] += hw.tree[1].proxyCount + hw.tree[2].proxyCount;
    for (int i = 0; i < hw.contactCount && i < (int)hw.contacts.size(); ++i) {
      const Contact& c = hw.contacts[i];
      if (c.shapeA < 0) continue;
      v[9] += 1;
      if ((c.flags & CF_TOUCHING) == 0) continue;
      v[2] += 1;
      v[3] += c.m.pointCount;
      if (c.restitution > 0.0f) {
        v[4] += 1;
        v[5] += c.m.pointCount;
      }
    }
    v[6] += (long)hw.joints.size();
    v[7] += (long)hw.overflow.size();
  }
  for (int k = 0; k < 10; ++k) out[k] = v[k];
  return 0;
}

// ---- flat per-world ABI (F/native/wasm/f2d_shim.c) ---------------------------

// f2d_shim.c:123-139
uint32_t f2d_create_world(float gravity_x, float gravity_y, float restitution_threshold, float hit_event_threshold,
                          float contact_hertz, float contact_damping_ratio, float max_contact_push_speed,
                          float maximum_linear_speed, int enable_sleep, int enable_continuous) {
  std::lock_guard<std::mutex> g(g_lock);
  Batch* b = new Batch();
  b->worlds.resize(1);
  SpecWorld sw = {};
  sw.flags = (enable_sleep ? 1 : 0) | (enable_continuous ? 2 : 0) | 4;
  sw.gx = gravity_x;
  sw.gy = gravity_y;
  sw.restitution_threshold = restitution_threshold;
  sw.hit_event_threshold = hit_event_threshold;
  sw.contact_hertz = contact_hertz;
  sw.contact_damping_ratio = contact_damping_ratio;
  sw.max_contact_push_speed = max_contact_push_speed;
  sw.maximum_linear_speed = maximum_linear_speed;
  initWorldHeader(b->worlds[0], sw);
  size_t slot = g_worlds.size();
  for (size_t i = 0; i < g_worlds.size(); ++i) {
    if (g_worlds[i].batch == nullptr) {
      slot = i;
      break;
    }
  }
  if (slot == g_worlds.size()) g_worlds.push_back(WorldSlot{});
  g_worlds[slot].batch = b;
  g_worlds[slot].generation += 1;
  return (uint32_t)(slot + 1) | ((uint32_t)g_worlds[slot].generation << 16);
}

void f2d_destroy_world(uint32_t w) {
  std::lock_guard<std::mutex> g(g_lock);
  Batch* b = batchOf(w);
  if (!b) return;
  f2dg_destroy(b);
  g_worlds[(w & 0xFFFF) - 1].batch = nullptr;
}

int f2d_world_is_valid(uint32_t w) { return batchOf(w) != nullptr; }

// f2d_shim.c:147-149 -> b2World_Step
void f2d_world_step(uint32_t w, float time_step, int sub_steps) {
  Batch* b = batchOf(w);
  if (!b) return;
  if (time_step == 0.0f) {
    // a paused step only retires the last step's events (S/world.c:709-725)
    download(*b);
    for (HostWorld& hw : b->worlds) {
      hw.beginIds.clear();
      hw.endPairs.clear();
      hw.hdr.beginCount = hw.hdr.endCount = 0;
      hw.hdr.moveBodyCount = 0;
      hw.eventsRetired = true;
    }
    return;
  }
  // a failed launch or a device fault ends the process, as the reference's asserts do
  if (f2dg_step(b, time_step, sub_steps, 1, PH_ALL) != 0) {
    hipError_t e = hipGetLastError();
    std::fprintf(stderr, "forge2d_amd: world step launch failed: %s (%s)\n", hipGetErrorString(e),
                 b->lastError.c_str());
    std::abort();
  }
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "forge2d_amd: world step failed on the device: %s\n", hipGetErrorString(e));
    std::abort();
  }
  const int bits = f2dg_error_bits(b);
  if (bits != 0) {
    std::fprintf(stderr, "forge2d_amd: world step exceeded a device capacity (error bits 0x%x:%s%s%s%s%s)\n", bits,
                 (bits & ERR_CONTACT_CAP) ? " contacts" : "", (bits & ERR_PAIR_CAP) ? " new pairs" : "",
                 (bits & ERR_OVERFLOW_CAP) ? " overflow constraints" : "", (bits & ERR_SCRATCH) ? " scratch" : "",
                 (bits & ERR_TREE) ? " tree" : "");
    std::abort();
  }
}

void f2d_world_set_gravity(uint32_t w, float x, float y) {
  Batch* b = batchOf(w);
  if (!b) return;
  download(*b);
  b->worlds[0].hdr.gravity = {x, y};
  b->hostDirty = true;
}

void f2d_world_get_gravity(uint32_t w, float* out) {
  Batch* b = batchOf(w);
  if (!b) return;
  out[0] = b->worlds[0].hdr.gravity.x;
  out[1] = b->worlds[0].hdr.gravity.y;
}

// Island operations of the body and world API run on the device copy of the world
// (the wake and sleep code of the step): the host's edits are uploaded first and the
// mirror is refreshed on the next read.
static void islandCommand(Batch& b, int world, int op, int arg) {
  if (b.hostDirty || !b.committed) {
    if (upload(b) != 0) return;
  }
  if (f2d_launch_island_command(&b.dev.view, world, op, arg, b.stream) != 0) return;
  (void)hipStreamSynchronize(b.stream);
  b.deviceNewer = true;
}

// b2World_EnableSleeping, S/world.c:1710-1737: disabling sleep wakes every sleeping set
void f2d_world_enable_sleeping(uint32_t w, int enabled) {
  Batch* b = batchOf(w);
  if (!b) return;
  download(*b);
  WorldHeader& h = b->worlds[0].hdr;
  const bool was = (h.flags & WF_SLEEP) != 0;
  if (was == (enabled != 0)) return;
  h.flags = enabled ? (h.flags | WF_SLEEP) : (h.flags & ~WF_SLEEP);
  b->hostDirty = true;
  if (!enabled) islandCommand(*b, 0, 2, 0);
}

int f2d_world_is_sleeping_enabled(uint32_t w) {
  Batch* b = batchOf(w);
  return b ? ((b->worlds[0].hdr.flags & WF_SLEEP) != 0) : 0;
}

void f2d_world_enable_continuous(uint32_t w, int enabled) {
  Batch* b = batchOf(w);
  if (!b) return;
  download(*b);
  WorldHeader& h = b->worlds[0].hdr;
  h.flags = enabled ? (h.flags | WF_CONTINUOUS) : (h.flags & ~WF_CONTINUOUS);
  b->hostDirty = true;
}

int f2d_world_is_continuous_enabled(uint32_t w) {
  Batch* b = batchOf(w);
  return b ? ((b->worlds[0].hdr.flags & WF_CONTINUOUS) != 0) : 0;
}

// f2d_shim.c:192-218
void f2d_create_body(uint32_t w, int type, float px, float py, float qc, float qs, float lvx, float lvy,
                     float angular_velocity, float linear_damping, float angular_damping, float gravity_scale,
                     float sleep_threshold, const char* name, int enable_sleep, int is_awake, int fixed_rotation,
                     int is_bullet, int is_enabled, int allow_fast_rotation, int32_t* out) {
  (void)name;
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  if (!b) return;
  if (!is_enabled) {
    // the reference puts the body in the disabled set (no proxies, no simulation until
    // enabled); that set is not kept here: refused with a null id and an error
    b->lastError = "f2d_create_body: bodies created disabled (is_enabled = 0) are not supported";
    std::fprintf(stderr, "forge2d_amd: %s\n", b->lastError.c_str());
    return;
  }
  download(*b);
  if (b->committed) freeDevice(*b);  // new bodies change the batch layout
  int flags = 0;
  if (is_bullet) flags |= BF_BULLET;
  if (fixed_rotation) flags |= BF_FIXED_ROTATION;
  if (allow_fast_rotation) flags |= BF_ALLOW_FAST_ROTATION;
  if (enable_sleep) flags |= BF_ENABLE_SLEEP;
  HostWorld& hw = b->worlds[0];
  int id = hw.createBody(type, {px, py}, {qc, qs}, {lvx, lvy}, angular_velocity, linear_damping, angular_damping,
                         gravity_scale, sleep_threshold, flags, is_awake != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw.bodyGen[id] << 16));
}

int f2d_body_is_valid(int32_t index1, uint32_t wg) {
  int id;
  return bodyWorld(index1, wg, &id) != nullptr;
}

void f2d_body_get_position(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->sims[id].xf.p.x;
  out[1] = hw->sims[id].xf.p.y;
}

void f2d_body_get_rotation(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->sims[id].xf.q.c;
  out[1] = hw->sims[id].xf.q.s;
}

void f2d_body_get_linear_velocity(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->states[id].v.x;
  out[1] = hw->states[id].v.y;
}

float f2d_body_get_angular_velocity(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->states[id].w : 0.0f;
}

// b2WakeBody, S/body.c:332-344
static void wakeBody(uint32_t wg, int id) {
  Batch* b = bodyBatch(wg);
  if (b && b->worlds[0].sims[id].setIndex >= kFirstSleepingSet) islandCommand(*b, 0, 1, id);
}

// b2Body_SetLinearVelocity, S/body.c:767-788: a non-zero velocity wakes the body;
// a body that stays asleep keeps no velocity
void f2d_body_set_linear_velocity(int32_t index1, uint32_t wg, float x, float y) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw || hw->sims[id].type == kStatic) return;
  if (x * x + y * y > 0.0f) {
    wakeBody(wg, id);
    hw = bodyWorld(index1, wg, &id);
  }
  if (hw->sims[id].setIndex != kAwakeSet) return;
  hw->states[id].v = {x, y};
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_ApplyForceToCenter, S/body.c:918-933: wakes on request; forces reach awake bodies only
void f2d_body_apply_force_to_center(int32_t index1, uint32_t wg, float fx, float fy, int wake) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if (wake && hw->sims[id].setIndex >= kFirstSleepingSet) {
    wakeBody(wg, id);
    hw = bodyWorld(index1, wg, &id);
  }
  if (!hw || hw->sims[id].setIndex != kAwakeSet) return;
  hw->sims[id].force = add(hw->sims[id].force, {fx, fy});
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_ApplyLinearImpulseToCenter, S/body.c:975-995: v += invMass * impulse, then
// b2LimitVelocity (S/body.c:28-35)
void f2d_body_apply_linear_impulse_to_center(int32_t index1, uint32_t wg, float ix, float iy, int wake) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if (wake && hw->sims[id].setIndex >= kFirstSleepingSet) {
    wakeBody(wg, id);
    hw = bodyWorld(index1, wg, &id);
  }
  if (!hw || hw->sims[id].setIndex != kAwakeSet) return;
  BodyState& st = hw->states[id];
  st.v = mulAdd(st.v, hw->sims[id].invMass, {ix, iy});
  const float v2 = lengthSq(st.v), maxSpeed = hw->hdr.maxLinearSpeed;
  if (v2 > maxSpeed * maxSpeed) st.v = mulSV(maxSpeed / sqrtf(v2), st.v);
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_IsAwake, S/body.c:1476-1481
int f2d_body_is_awake(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? (hw->sims[id].setIndex == kAwakeSet ? 1 : 0) : 0;
}

// b2Body_SetAwake, S/body.c:1483-1515
void f2d_body_set_awake(int32_t index1, uint32_t wg, int awake) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if (awake && hw->sims[id].setIndex >= kFirstSleepingSet) {
    wakeBody(wg, id);
  } else if (!awake && hw->sims[id].setIndex == kAwakeSet) {
    islandCommand(*bodyBatch(wg), 0, 3, id);
  }
}

// b2Body_IsSleepEnabled / b2Body_EnableSleep, S/body.c:1517-1522, 1538-1552: disabling sleep wakes the body
int f2d_body_is_sleep_enabled(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? ((hw->sims[id].flags & BF_ENABLE_SLEEP) ? 1 : 0) : 0;
}

void f2d_body_enable_sleep(int32_t index1, uint32_t wg, int enable) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if (enable) {
    hw->sims[id].flags |= BF_ENABLE_SLEEP;
  } else {
    hw->sims[id].flags &= ~BF_ENABLE_SLEEP;
  }
  bodyBatch(wg)->hostDirty = true;
  if (!enable) wakeBody(wg, id);
}

// b2Body_GetSleepThreshold / b2Body_SetSleepThreshold, S/body.c:1524-1536
float f2d_body_get_sleep_threshold(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].sleepThreshold : 0.0f;
}

void f2d_body_set_sleep_threshold(int32_t index1, uint32_t wg, float threshold) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  hw->sims[id].sleepThreshold = threshold;
  bodyBatch(wg)->hostDirty = true;
}

void f2d_body_get_world_center(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->sims[id].center.x;
  out[1] = hw->sims[id].center.y;
}

void f2d_body_get_local_center(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->sims[id].localCenter.x;
  out[1] = hw->sims[id].localCenter.y;
}

// ---- more of the body API (S/body.c), on the host mirror ---------------------

// b2Body_SetTransform, S/body.c:688-735: the shapes' boxes follow the new transform;
// a fat box that no longer holds its shape's box grows and its proxy moves in its
// tree (b2BroadPhase_MoveProxy, S/broad_phase.c:116-123: removed, reinserted without
// rotations, and buffered as moved)
void f2d_body_set_transform(int32_t index1, uint32_t wg, float px, float py, float qc, float qs) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  BodySim& sim = hw->sims[id];
  sim.xf.p = {px, py};
  sim.xf.q = {qc, qs};
  sim.center = transformPoint(sim.xf, sim.localCenter);
  sim.rotation0 = sim.xf.q;
  sim.center0 = sim.center;
  const Xf xf = sim.xf;
  for (int s = sim.headShape, g = 0; s >= 0 && g < (int)hw->shapes.size(); s = hw->shapes[s].next, ++g) {
    Shape& shape = hw->shapes[s];
    AABB aabb = computeShapeAABB(hw->geoms[s], xf);
    aabb.lo.x -= kSpeculativeDistance;
    aabb.lo.y -= kSpeculativeDistance;
    aabb.hi.x += kSpeculativeDistance;
    aabb.hi.y += kSpeculativeDistance;
    shape.aabb = aabb;
    if (!aabbContains(shape.fat, aabb)) {
      AABB fat;
      fat.lo.x = aabb.lo.x - kAabbMargin;
      fat.lo.y = aabb.lo.y - kAabbMargin;
      fat.hi.x = aabb.hi.x + kAabbMargin;
      fat.hi.y = aabb.hi.y + kAabbMargin;
      shape.fat = fat;
      if (shape.proxyKey >= 0) {
        shape.flags &= ~SF_TREE_PENDING;
        treeMoveProxy(hw->treeRef(shape.proxyKey & 3), shape.proxyKey >> 2, fat);
        if (!(shape.flags & SF_MOVED)) {  // b2BufferMove: a proxy enters the move array once
          shape.flags |= SF_MOVED;
          shape.moveRank = (1 << 30) + hw->createdSeq++;
        }
      }
    }
  }
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_SetAngularVelocity, S/body.c:791-812
void f2d_body_set_angular_velocity(int32_t index1, uint32_t wg, float w) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw || hw->sims[id].type == kStatic || hw->bodyFixedRotation[id]) return;
  if (w != 0.0f) {
    wakeBody(wg, id);
    hw = bodyWorld(index1, wg, &id);
  }
  if (hw->sims[id].setIndex != kAwakeSet) return;
  hw->states[id].w = w;
  bodyBatch(wg)->hostDirty = true;
}

// the awake body's sim and state after an optional wake (the force and impulse calls)
static HostWorld* awakeTarget(int32_t index1, uint32_t wg, int wake, int* id) {
  HostWorld* hw = bodyWorld(index1, wg, id);
  if (!hw) return nullptr;
  if (wake && hw->sims[*id].setIndex >= kFirstSleepingSet) {
    wakeBody(wg, *id);
    hw = bodyWorld(index1, wg, id);
  }
  if (!hw || hw->sims[*id].setIndex != kAwakeSet) return nullptr;
  bodyBatch(wg)->hostDirty = true;
  return hw;
}

// b2LimitVelocity, S/body.c:28-35
static void limitVelocity(BodyState& st, float maxLinearSpeed) {
  const float v2 = lengthSq(st.v);
  if (v2 > maxLinearSpeed * maxLinearSpeed) st.v = mulSV(maxLinearSpeed / sqrtf(v2), st.v);
}

// b2Body_ApplyForce, S/body.c:900-916
void f2d_body_apply_force(int32_t index1, uint32_t wg, float fx, float fy, float px, float py, int wake) {
  int id;
  HostWorld* hw = awakeTarget(index1, wg, wake, &id);
  if (!hw) return;
  BodySim& sim = hw->sims[id];
  sim.force = add(sim.force, {fx, fy});
  sim.torque += cross(sub({px, py}, sim.center), {fx, fy});
}

// b2Body_ApplyTorque, S/body.c:935-950
void f2d_body_apply_torque(int32_t index1, uint32_t wg, float torque, int wake) {
  int id;
  HostWorld* hw = awakeTarget(index1, wg, wake, &id);
  if (!hw) return;
  hw->sims[id].torque += torque;
}

// b2Body_ApplyLinearImpulse, S/body.c:952-973
void f2d_body_apply_linear_impulse(int32_t index1, uint32_t wg, float ix, float iy, float px, float py, int wake) {
  int id;
  HostWorld* hw = awakeTarget(index1, wg, wake, &id);
  if (!hw) return;
  const BodySim& sim = hw->sims[id];
  BodyState& st = hw->states[id];
  st.v = mulAdd(st.v, sim.invMass, {ix, iy});
  st.w += sim.invI * cross(sub({px, py}, sim.center), {ix, iy});
  limitVelocity(st, hw->hdr.maxLinearSpeed);
}

// b2Body_ApplyAngularImpulse, S/body.c:997-1020
void f2d_body_apply_angular_impulse(int32_t index1, uint32_t wg, float impulse, int wake) {
  int id;
  HostWorld* hw = awakeTarget(index1, wg, wake, &id);
  if (!hw) return;
  hw->states[id].w += hw->sims[id].invI * impulse;
}

// b2Body_SetMassData, S/body.c:1357-1381
void f2d_body_set_mass_data(int32_t index1, uint32_t wg, float mass, float rotational_inertia, float cx, float cy) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  BodySim& sim = hw->sims[id];
  hw->bodyMass[id] = mass;
  hw->bodyInertia[id] = rotational_inertia;
  sim.localCenter = {cx, cy};
  const Vec2 center = transformPoint(sim.xf, sim.localCenter);
  sim.center = center;
  sim.center0 = center;
  sim.invMass = mass > 0.0f ? 1.0f / mass : 0.0f;
  sim.invI = rotational_inertia > 0.0f ? 1.0f / rotational_inertia : 0.0f;
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_ApplyMassFromShapes, S/body.c:1393-1403 (b2UpdateBodyMassData)
void f2d_body_apply_mass_from_shapes(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  hw->updateBodyMass(id);
  bodyBatch(wg)->hostDirty = true;
}

// b2Body_GetWorldPoint / GetLocalPoint, S/body.c:656-670
void f2d_body_get_world_point(int32_t index1, uint32_t wg, float x, float y, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  const Vec2 p = transformPoint(hw->sims[id].xf, {x, y});
  out[0] = p.x;
  out[1] = p.y;
}

void f2d_body_get_local_point(int32_t index1, uint32_t wg, float x, float y, float* out) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  const Vec2 p = invTransformPoint(hw->sims[id].xf, {x, y});
  out[0] = p.x;
  out[1] = p.y;
}

// b2Body_GetShapeCount / GetShapes, S/body.c:1797-1827: the body's shape list, newest first
int f2d_body_get_shape_count(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return 0;
  int n = 0;
  for (int s = hw->sims[id].headShape; s >= 0 && n <= (int)hw->shapes.size(); s = hw->shapes[s].next) ++n;
  return n;
}

int f2d_body_get_shapes(int32_t index1, uint32_t wg, int32_t* out, int capacity) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return 0;
  int n = 0;
  for (int s = hw->sims[id].headShape; s >= 0 && n < capacity; s = hw->shapes[s].next) {
    out[2 * n] = s + 1;
    out[2 * n + 1] = (int32_t)((wg & 0xFFFF) | (1u << 16));
    ++n;
  }
  return n;
}

// b2Body_GetJointCount / GetJoints, S/body.c:1829-1856: the body's joint list by edge keys
int f2d_body_get_joint_count(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].jointCount : 0;
}

int f2d_body_get_joints(int32_t index1, uint32_t wg, int32_t* out, int capacity) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return 0;
  int n = 0;
  for (int key = hw->sims[id].headJointKey; key >= 0 && n < capacity;) {
    const int j = key >> 1;
    out[2 * n] = j + 1;
    out[2 * n + 1] = (int32_t)((wg & 0xFFFF) | ((uint32_t)hw->jointGen[j] << 16));
    ++n;
    key = (key & 1) ? hw->joints[j].nextB : hw->joints[j].nextA;
  }
  return n;
}

// b2Body_GetName / SetName, S/body.c (31 characters kept)
const char* f2d_body_get_name(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return "";
  if ((int)hw->bodyNames.size() <= id) hw->bodyNames.resize(hw->sims.size());
  return hw->bodyNames[id].c_str();
}

void f2d_body_set_name(int32_t index1, uint32_t wg, const char* name) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if ((int)hw->bodyNames.size() <= id) hw->bodyNames.resize(hw->sims.size());
  hw->bodyNames[id] = name ? std::string(name).substr(0, 31) : std::string();
}

// Scalar body accessors of the shim (S/body.c): gravity scale, damping, bullet, fixed
// rotation; every body here is enabled (disabled bodies are not supported)
int f2d_body_is_enabled(int32_t index1, uint32_t wg) {
  int id;
  return bodyWorld(index1, wg, &id) ? 1 : 0;
}

float f2d_body_get_gravity_scale(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].gravityScale : 0.0f;
}

void f2d_body_set_gravity_scale(int32_t index1, uint32_t wg, float v) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  hw->sims[id].gravityScale = v;
  bodyBatch(wg)->hostDirty = true;
}

float f2d_body_get_linear_damping(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].linearDamping : 0.0f;
}

void f2d_body_set_linear_damping(int32_t index1, uint32_t wg, float v) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  hw->sims[id].linearDamping = v;
  bodyBatch(wg)->hostDirty = true;
}

float f2d_body_get_angular_damping(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].angularDamping : 0.0f;
}

void f2d_body_set_angular_damping(int32_t index1, uint32_t wg, float v) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  hw->sims[id].angularDamping = v;
  bodyBatch(wg)->hostDirty = true;
}

int f2d_body_is_bullet(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? ((hw->sims[id].flags & BF_BULLET) ? 1 : 0) : 0;
}

// b2Body_SetBullet, S/body.c:1751-1762
void f2d_body_set_bullet(int32_t index1, uint32_t wg, int flag) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  if (flag) {
    hw->sims[id].flags |= BF_BULLET;
  } else {
    hw->sims[id].flags &= ~BF_BULLET;
  }
  bodyBatch(wg)->hostDirty = true;
}

int f2d_body_is_fixed_rotation(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->bodyFixedRotation[id] : 0;
}

// b2Body_SetFixedRotation, S/body.c:1722-1742: a change zeroes an awake body's angular
// velocity and recomputes its mass data
void f2d_body_set_fixed_rotation(int32_t index1, uint32_t wg, int flag) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw || (hw->bodyFixedRotation[id] != 0) == (flag != 0)) return;
  hw->bodyFixedRotation[id] = flag ? 1 : 0;
  if (flag) {
    hw->sims[id].flags |= BF_FIXED_ROTATION;
  } else {
    hw->sims[id].flags &= ~BF_FIXED_ROTATION;
  }
  if (hw->sims[id].setIndex == kAwakeSet) hw->states[id].w = 0.0f;
  hw->updateBodyMass(id);
  bodyBatch(wg)->hostDirty = true;
}

// ---- shape accessors (S/shape.c) ---------------------------------------------

static HostWorld* shapeWorld(int32_t index1, uint32_t wg, int* shapeId) {
  uint32_t world0 = wg & 0xFFFF;
  if (world0 >= g_worlds.size() || (wg >> 16) != 1u) return nullptr;
  WorldSlot& s = g_worlds[world0];
  if (!s.batch) return nullptr;
  if (s.batch->deviceNewer && download(*s.batch) != 0) return nullptr;
  HostWorld& hw = s.batch->worlds[0];
  const int id = index1 - 1;
  if (id < 0 || id >= (int)hw.shapes.size()) return nullptr;
  *shapeId = id;
  return &hw;
}

int f2d_shape_is_valid(int32_t index1, uint32_t wg) {
  int id;
  return shapeWorld(index1, wg, &id) ? 1 : 0;
}

void f2d_shape_get_body(int32_t index1, uint32_t wg, int32_t* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  const int b = hw->shapes[id].body;
  out[0] = b + 1;
  out[1] = (int32_t)((wg & 0xFFFF) | ((uint32_t)hw->bodyGen[b] << 16));
}

// shape type in the reference's b2ShapeType numbering (circle 0, capsule 1, segment 2, polygon 3)
int f2d_shape_get_type(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  return hw ? hw->geoms[id].type : -1;
}

int f2d_shape_is_sensor(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  return hw ? ((hw->shapes[id].flags & SF_SENSOR) ? 1 : 0) : 0;
}

float f2d_shape_get_density(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  return hw ? hw->shapeDensity[id] : 0.0f;
}

// b2Shape_SetDensity, S/shape.c:1055-1086
void f2d_shape_set_density(int32_t index1, uint32_t wg, float density, int update_body_mass) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw || density == hw->shapeDensity[id]) return;
  hw->shapeDensity[id] = density;
  if (update_body_mass) hw->updateBodyMass(hw->shapes[id].body);
  g_worlds[wg & 0xFFFF].batch->hostDirty = true;
}

float f2d_shape_get_friction(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  return hw ? hw->shapes[id].friction : 0.0f;
}

// b2Shape_SetFriction / SetRestitution, S/shape.c:1088-1130: touching contacts take the
// new values at their next update
void f2d_shape_set_friction(int32_t index1, uint32_t wg, float v) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  hw->shapes[id].friction = v;
  g_worlds[wg & 0xFFFF].batch->hostDirty = true;
}

float f2d_shape_get_restitution(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  return hw ? hw->shapes[id].restitution : 0.0f;
}

void f2d_shape_set_restitution(int32_t index1, uint32_t wg, float v) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  hw->shapes[id].restitution = v;
  g_worlds[wg & 0xFFFF].batch->hostDirty = true;
}

// b2Shape_GetFilter: [category lo, hi, mask lo, hi, group]
void f2d_shape_get_filter(int32_t index1, uint32_t wg, int32_t* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  const Shape& s = hw->shapes[id];
  out[0] = (int32_t)s.catLo;
  out[1] = (int32_t)s.catHi;
  out[2] = (int32_t)s.maskLo;
  out[3] = (int32_t)s.maskHi;
  out[4] = s.group;
}

// event flags of the shape (S/shape.c:1282-1340); sensor events are not kept
#define F2D_SHAPE_FLAG(getter, setter, bit)                                   \
  int getter(int32_t index1, uint32_t wg) {                                   \
    int id;                                                                   \
    HostWorld* hw = shapeWorld(index1, wg, &id);                              \
    return hw ? ((hw->shapes[id].flags & (bit)) ? 1 : 0) : 0;                 \
  }                                                                           \
  void setter(int32_t index1, uint32_t wg, int flag) {                        \
    int id;                                                                   \
    HostWorld* hw = shapeWorld(index1, wg, &id);                              \
    if (!hw) return;                                                          \
    if (flag) {                                                               \
      hw->shapes[id].flags |= (bit);                                          \
    } else {                                                                  \
      hw->shapes[id].flags &= ~(bit);                                         \
    }                                                                         \
    g_worlds[wg & 0xFFFF].batch->hostDirty = true;                            \
  }
F2D_SHAPE_FLAG(f2d_shape_are_contact_events_enabled, f2d_shape_enable_contact_events, SF_CONTACT_EVENTS)
F2D_SHAPE_FLAG(f2d_shape_are_hit_events_enabled, f2d_shape_enable_hit_events, SF_HIT_EVENTS)
F2D_SHAPE_FLAG(f2d_shape_are_pre_solve_events_enabled, f2d_shape_enable_pre_solve_events, SF_PRESOLVE)
#undef F2D_SHAPE_FLAG

// b2Shape_TestPoint, S/shape.c:979-1000 with b2PointInCircle / Capsule / Polygon,
// S/geometry.c:456-501
int f2d_shape_test_point(int32_t index1, uint32_t wg, float x, float y) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return 0;
  const ShapeGeom& g = hw->geoms[id];
  const Vec2 p = invTransformPoint(hw->sims[hw->shapes[id].body].xf, {x, y});
  if (g.type == kCircle) return distanceSq(p, g.v[0]) <= g.radius * g.radius ? 1 : 0;
  if (g.type == kCapsule) {
    const float rr = g.radius * g.radius;
    const Vec2 p1 = g.v[0], p2 = g.v[1];
    const Vec2 d = sub(p2, p1);
    const float dd = dot(d, d);
    if (dd == 0.0f) return distanceSq(p, p1) <= rr ? 1 : 0;
    float t = dot(sub(p, p1), d) / dd;
    t = clampf(t, 0.0f, 1.0f);
    const Vec2 c = mulAdd(p1, t, d);
    return distanceSq(p, c) <= rr ? 1 : 0;
  }
  if (g.type == kPolygon) {
    Proxy pa;
    pa.count = g.count;
    for (int i = 0; i < g.count; ++i) pa.p[i] = g.v[i];
    pa.radius = 0.0f;
    Proxy pb;
    pb.count = 1;
    pb.p[0] = p;
    pb.radius = 0.0f;
    SimplexCache cache;
    cache.count = 0;
    const Xf identity = {{0.0f, 0.0f}, {1.0f, 0.0f}};
    const DistanceOut o = shapeDistanceOut(cache, pa, pb, identity, identity, false);
    return o.distance <= g.radius ? 1 : 0;
  }
  return 0;
}

void f2d_shape_get_circle(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->geoms[id].v[0].x;
  out[1] = hw->geoms[id].v[0].y;
  out[2] = hw->geoms[id].radius;
}

void f2d_shape_get_capsule(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  const ShapeGeom& g = hw->geoms[id];
  out[0] = g.v[0].x;
  out[1] = g.v[0].y;
  out[2] = g.v[1].x;
  out[3] = g.v[1].y;
  out[4] = g.radius;
}

void f2d_shape_get_segment(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  const ShapeGeom& g = hw->geoms[id];
  out[0] = g.v[0].x;
  out[1] = g.v[0].y;
  out[2] = g.v[1].x;
  out[3] = g.v[1].y;
}

// the radius, then the vertex pairs; returns the vertex count
int f2d_shape_get_polygon(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return 0;
  const ShapeGeom& g = hw->geoms[id];
  out[0] = g.radius;
  for (int i = 0; i < g.count; ++i) {
    out[1 + 2 * i] = g.v[i].x;
    out[2 + 2 * i] = g.v[i].y;
  }
  return g.count;
}

// b2Shape_GetAABB: the shape's box as the last step (or call) left it
void f2d_shape_get_aabb(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = shapeWorld(index1, wg, &id);
  if (!hw) return;
  const AABB a = hw->shapes[id].aabb;
  out[0] = a.lo.x;
  out[1] = a.lo.y;
  out[2] = a.hi.x;
  out[3] = a.hi.y;
}

// ---- joint accessors (S/joint.c, S/revolute_joint.c) -------------------------

static HostWorld* jointWorld(int32_t index1, uint32_t wg, int* jointId) {
  uint32_t world0 = wg & 0xFFFF;
  if (world0 >= g_worlds.size()) return nullptr;
  WorldSlot& s = g_worlds[world0];
  if (!s.batch) return nullptr;
  if (s.batch->deviceNewer && download(*s.batch) != 0) return nullptr;
  HostWorld& hw = s.batch->worlds[0];
  const int id = index1 - 1;
  if (id < 0 || id >= (int)hw.joints.size()) return nullptr;
  if (hw.joints[id].setIndex < 0 || (id < (int)hw.jointGen.size() && (wg >> 16) != hw.jointGen[id])) return nullptr;
  *jointId = id;
  return &hw;
}

// f2d_shim.c:890-893 (b2DestroyJoint, S/joint.c:807-819). A joint of a sleeping or the
// static set is left alone with a message (its set arrays are not restated).
void f2d_destroy_joint(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  if (hw->destroyJoint(id) != 0) {
    fprintf(stderr, "forge2d_amd: destroying joint %d outside the awake set is not supported\n", id);
    return;
  }
  bodyBatch(wg)->hostDirty = true;
}

static bool refuseContactsBetween(const HostWorld& hw, int a, int b, int collideConnected);

// f2d_shim.c:932-935 (b2Joint_SetCollideConnected, S/joint.c:979-1022): turned on, the
// shapes of the body with fewer shapes (body B on a tie) enter the move buffer so the
// next step looks for their pairs; turned off, the contacts between the bodies are
// destroyed in the reference, which is not restated: refused while they share one.
void f2d_joint_set_collide_connected(int32_t index1, uint32_t wg, int v) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  Joint& j = hw->joints[id];
  const int should = v != 0 ? 1 : 0;
  if (j.collideConnected == should) return;
  if (!should) {
    if (refuseContactsBetween(*hw, j.bodyA, j.bodyB, 0)) return;
    j.collideConnected = 0;
    bodyBatch(wg)->hostDirty = true;
    return;
  }
  j.collideConnected = 1;
  auto shapeCount = [&](int body) {
    int n = 0;
    for (int s = hw->sims[body].headShape, g = 0; s >= 0 && g < (int)hw->shapes.size(); s = hw->shapes[s].next, ++g) ++n;
    return n;
  };
  const int body = shapeCount(j.bodyA) < shapeCount(j.bodyB) ? j.bodyA : j.bodyB;
  for (int s = hw->sims[body].headShape, g = 0; s >= 0 && g < (int)hw->shapes.size(); s = hw->shapes[s].next, ++g) {
    Shape& shape = hw->shapes[s];
    if (shape.proxyKey >= 0 && !(shape.flags & SF_MOVED)) {  // b2BufferMove
      shape.flags |= SF_MOVED;
      shape.moveRank = (1 << 30) + hw->createdSeq++;
    }
  }
  bodyBatch(wg)->hostDirty = true;
}

int f2d_joint_is_valid(int32_t index1, uint32_t wg) {
  int id;
  return jointWorld(index1, wg, &id) ? 1 : 0;
}

// b2JointType numbering (I/types.h): every joint here is revolute (5)
int f2d_joint_get_type(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  return hw ? hw->joints[id].type : -1;
}

static void jointBody(const HostWorld& hw, int body, uint32_t wg, int32_t* out) {
  out[0] = body + 1;
  out[1] = (int32_t)((wg & 0xFFFF) | ((uint32_t)hw.bodyGen[body] << 16));
}

void f2d_joint_get_body_a(int32_t index1, uint32_t wg, int32_t* out) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (hw) jointBody(*hw, hw->joints[id].bodyA, wg, out);
}

void f2d_joint_get_body_b(int32_t index1, uint32_t wg, int32_t* out) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (hw) jointBody(*hw, hw->joints[id].bodyB, wg, out);
}

// b2Joint_GetLocalAnchorA / B: the anchors in the bodies' origin frames
void f2d_joint_get_local_anchor_a(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->joints[id].localOriginAnchorA.x;
  out[1] = hw->joints[id].localOriginAnchorA.y;
}

void f2d_joint_get_local_anchor_b(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  out[0] = hw->joints[id].localOriginAnchorB.x;
  out[1] = hw->joints[id].localOriginAnchorB.y;
}

int f2d_joint_get_collide_connected(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  return hw ? hw->joints[id].collideConnected : 0;
}

// b2Joint_WakeBodies, S/joint.c: both bodies' sleeping sets wake
void f2d_joint_wake_bodies(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  const int a = hw->joints[id].bodyA, b = hw->joints[id].bodyB;
  const uint32_t bwg = (wg & 0xFFFF);
  wakeBody(bwg | ((uint32_t)hw->bodyGen[a] << 16), a);
  hw = jointWorld(index1, wg, &id);
  if (hw) wakeBody(bwg | ((uint32_t)hw->bodyGen[b] << 16), b);
}

// b2GetRevoluteJointForce / Torque, S/revolute_joint.c:176-187 (world inv_h of the last step)
// b2Joint_GetConstraintForce: revolute S/revolute_joint.c (inv_h x linear impulse),
// prismatic S/prismatic_joint.c:211-227 (perpendicular and axial parts in body A's
// current frame)
void f2d_joint_get_constraint_force(int32_t index1, uint32_t wg, float* out) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return;
  const Joint& j = hw->joints[id];
  Vec2 f;
  if (j.type == kPrismaticJoint) {
    const Vec2 axisA = rotate(hw->sims[j.bodyA].xf.q, j.localAxisA);
    const Vec2 perpA = leftPerp(axisA);
    const float perpForce = hw->inv_h * j.linearImpulse.x;
    const float axialForce = hw->inv_h * (j.motorImpulse + j.lowerImpulse - j.upperImpulse);
    f = add(mulSV(perpForce, perpA), mulSV(axialForce, axisA));
  } else if (j.type == kWheelJoint) {
    // S/wheel_joint.c:147-160: the axis of the last prepare (a frame behind)
    const Vec2 perpA = leftPerp(j.axisA);
    const float perpForce = hw->inv_h * j.linearImpulse.x;
    const float axialForce = hw->inv_h * (j.springImpulse + j.lowerImpulse - j.upperImpulse);
    f = add(mulSV(perpForce, perpA), mulSV(axialForce, j.axisA));
  } else if (j.type == kDistanceJoint) {
    // S/distance_joint.c:185-198: along the current anchor-to-anchor axis
    const Vec2 pA = transformPoint(hw->sims[j.bodyA].xf, j.localOriginAnchorA);
    const Vec2 pB = transformPoint(hw->sims[j.bodyB].xf, j.localOriginAnchorB);
    const Vec2 axis = normalize(sub(pB, pA));
    const float force = (j.linearImpulse.x + j.lowerImpulse - j.upperImpulse + j.motorImpulse) * hw->inv_h;
    f = mulSV(force, axis);
  } else if (j.type == kFilterJoint) {
    f = {0.0f, 0.0f};
  } else {
    // revolute, weld (S/weld_joint.c:66-70), motor (S/motor_joint.c:73-77), mouse
    // (S/mouse_joint.c:65-69)
    f = mulSV(hw->inv_h, j.linearImpulse);
  }
  out[0] = f.x;
  out[1] = f.y;
}

// b2Joint_GetConstraintTorque: prismatic S/prismatic_joint.c:229-232
float f2d_joint_get_constraint_torque(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw) return 0.0f;
  const Joint& j = hw->joints[id];
  if (j.type == kPrismaticJoint) return hw->inv_h * j.linearImpulse.y;
  if (j.type == kWeldJoint) return hw->inv_h * j.springImpulse;  // angular impulse, S/weld_joint.c:72-75
  if (j.type == kWheelJoint) return hw->inv_h * j.motorImpulse;  // S/wheel_joint.c:162-165
  if (j.type == kDistanceJoint || j.type == kFilterJoint) return 0.0f;  // S/joint.c:1100-1120
  if (j.type == kMotorJoint || j.type == kMouseJoint) return hw->inv_h * j.springImpulse;  // angular impulse
  return hw->inv_h * (j.motorImpulse + j.lowerImpulse - j.upperImpulse);
}

// ---- weld and wheel joint parameters (f2d_shim.c:1138-1185 -> S/weld_joint.c:14-64,
// S/wheel_joint.c:12-145). A call on a joint of another type is ignored (getters return 0).
static Joint* jointOfType(int32_t index1, uint32_t wg, int type, HostWorld** world = nullptr) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw || hw->joints[id].type != type) return nullptr;
  if (world) *world = hw;
  return &hw->joints[id];
}
#define F2D_JOINT_FLOAT(getter, setter, type, field)                      \
  float getter(int32_t index1, uint32_t wg) {                              \
    const Joint* j = jointOfType(index1, wg, type);                        \
    return j ? j->field : 0.0f;                                            \
  }                                                                        \
  void setter(int32_t index1, uint32_t wg, float v) {                      \
    Joint* j = jointOfType(index1, wg, type);                              \
    if (!j) return;                                                        \
    j->field = v;                                                          \
    bodyBatch(wg)->hostDirty = true;                                       \
  }
F2D_JOINT_FLOAT(f2d_weld_joint_get_linear_hertz, f2d_weld_joint_set_linear_hertz, kWeldJoint, hertz)
F2D_JOINT_FLOAT(f2d_weld_joint_get_angular_hertz, f2d_weld_joint_set_angular_hertz, kWeldJoint, angularHertz)
F2D_JOINT_FLOAT(f2d_weld_joint_get_linear_damping_ratio, f2d_weld_joint_set_linear_damping_ratio, kWeldJoint,
                dampingRatio)
F2D_JOINT_FLOAT(f2d_weld_joint_get_angular_damping_ratio, f2d_weld_joint_set_angular_damping_ratio, kWeldJoint,
                angularDampingRatio)
F2D_JOINT_FLOAT(f2d_wheel_joint_get_spring_hertz, f2d_wheel_joint_set_spring_hertz, kWheelJoint, hertz)
F2D_JOINT_FLOAT(f2d_wheel_joint_get_spring_damping_ratio, f2d_wheel_joint_set_spring_damping_ratio, kWheelJoint,
                dampingRatio)
F2D_JOINT_FLOAT(f2d_wheel_joint_get_motor_speed, f2d_wheel_joint_set_motor_speed, kWheelJoint, motorSpeed)
F2D_JOINT_FLOAT(f2d_wheel_joint_get_max_motor_torque, f2d_wheel_joint_set_max_motor_torque, kWheelJoint,
                maxMotorTorque)
F2D_JOINT_FLOAT(f2d_distance_joint_get_spring_hertz, f2d_distance_joint_set_spring_hertz, kDistanceJoint, hertz)
F2D_JOINT_FLOAT(f2d_distance_joint_get_spring_damping_ratio, f2d_distance_joint_set_spring_damping_ratio,
                kDistanceJoint, dampingRatio)
F2D_JOINT_FLOAT(f2d_distance_joint_get_motor_speed, f2d_distance_joint_set_motor_speed, kDistanceJoint, motorSpeed)
F2D_JOINT_FLOAT(f2d_distance_joint_get_max_motor_force, f2d_distance_joint_set_max_motor_force, kDistanceJoint,
                maxMotorTorque)

// ---- motor and mouse joint parameters (f2d_shim.c:1006-1050 -> S/motor_joint.c:14-71,
// S/mouse_joint.c:14-63): stored as given, no impulse reset and no wake; the motor's
// maximum force and torque are clamped at 0 and its correction factor to [0, 1]
F2D_JOINT_FLOAT(f2d_motor_joint_get_angular_offset, f2d_motor_joint_set_angular_offset, kMotorJoint, referenceAngle)
F2D_JOINT_FLOAT(f2d_mouse_joint_get_spring_hertz, f2d_mouse_joint_set_spring_hertz, kMouseJoint, hertz)
F2D_JOINT_FLOAT(f2d_mouse_joint_get_spring_damping_ratio, f2d_mouse_joint_set_spring_damping_ratio, kMouseJoint,
                dampingRatio)
F2D_JOINT_FLOAT(f2d_mouse_joint_get_max_force, f2d_mouse_joint_set_max_force, kMouseJoint, maxMotorTorque)
static float motorGet(int32_t index1, uint32_t wg, float Joint::*field) {
  const Joint* j = jointOfType(index1, wg, kMotorJoint);
  return j ? j->*field : 0.0f;
}
static void motorSet(int32_t index1, uint32_t wg, float Joint::*field, float v) {
  Joint* j = jointOfType(index1, wg, kMotorJoint);
  if (!j) return;
  j->*field = v;
  bodyBatch(wg)->hostDirty = true;
}
float f2d_motor_joint_get_max_force(int32_t index1, uint32_t wg) { return motorGet(index1, wg, &Joint::maxMotorTorque); }
void f2d_motor_joint_set_max_force(int32_t index1, uint32_t wg, float v) {
  motorSet(index1, wg, &Joint::maxMotorTorque, maxf(0.0f, v));
}
float f2d_motor_joint_get_max_torque(int32_t index1, uint32_t wg) { return motorGet(index1, wg, &Joint::angularHertz); }
void f2d_motor_joint_set_max_torque(int32_t index1, uint32_t wg, float v) {
  motorSet(index1, wg, &Joint::angularHertz, maxf(0.0f, v));
}
float f2d_motor_joint_get_correction_factor(int32_t index1, uint32_t wg) {
  return motorGet(index1, wg, &Joint::angularDampingRatio);
}
void f2d_motor_joint_set_correction_factor(int32_t index1, uint32_t wg, float v) {
  motorSet(index1, wg, &Joint::angularDampingRatio, clampf(v, 0.0f, 1.0f));
}
static void jointVec2(int32_t index1, uint32_t wg, int type, float* out) {
  const Joint* j = jointOfType(index1, wg, type);
  out[0] = j ? j->localAxisA.x : 0.0f;
  out[1] = j ? j->localAxisA.y : 0.0f;
}
static void setJointVec2(int32_t index1, uint32_t wg, int type, float x, float y) {
  Joint* j = jointOfType(index1, wg, type);
  if (!j) return;
  j->localAxisA = {x, y};
  bodyBatch(wg)->hostDirty = true;
}
void f2d_motor_joint_get_linear_offset(int32_t index1, uint32_t wg, float* out) {
  jointVec2(index1, wg, kMotorJoint, out);
}
void f2d_motor_joint_set_linear_offset(int32_t index1, uint32_t wg, float x, float y) {
  setJointVec2(index1, wg, kMotorJoint, x, y);
}
void f2d_mouse_joint_get_target(int32_t index1, uint32_t wg, float* out) { jointVec2(index1, wg, kMouseJoint, out); }
void f2d_mouse_joint_set_target(int32_t index1, uint32_t wg, float x, float y) {
  setJointVec2(index1, wg, kMouseJoint, x, y);
}

// ---- revolute joint (S/revolute_joint.c:20-176): the same rules as the prismatic calls
static void enableWithReset(Joint* j, int& flag, int v, float* a, float* b = nullptr) {
  if ((v != 0) != (flag != 0)) {
    flag = v != 0;
    *a = 0.0f;
    if (b) *b = 0.0f;
  }
}
F2D_JOINT_FLOAT(f2d_revolute_joint_get_spring_hertz, f2d_revolute_joint_set_spring_hertz, kRevoluteJoint, hertz)
F2D_JOINT_FLOAT(f2d_revolute_joint_get_spring_damping_ratio, f2d_revolute_joint_set_spring_damping_ratio,
                kRevoluteJoint, dampingRatio)
F2D_JOINT_FLOAT(f2d_revolute_joint_get_target_angle, f2d_revolute_joint_set_target_angle, kRevoluteJoint,
                targetAngle)
F2D_JOINT_FLOAT(f2d_revolute_joint_get_motor_speed, f2d_revolute_joint_set_motor_speed, kRevoluteJoint, motorSpeed)
F2D_JOINT_FLOAT(f2d_revolute_joint_get_max_motor_torque, f2d_revolute_joint_set_max_motor_torque, kRevoluteJoint,
                maxMotorTorque)
int f2d_revolute_joint_is_spring_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  return j ? j->enableSpring : 0;
}
void f2d_revolute_joint_enable_spring(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  if (!j) return;
  enableWithReset(j, j->enableSpring, v, &j->springImpulse);
  bodyBatch(wg)->hostDirty = true;
}
int f2d_revolute_joint_is_limit_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  return j ? j->enableLimit : 0;
}
void f2d_revolute_joint_enable_limit(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  if (!j) return;
  enableWithReset(j, j->enableLimit, v, &j->lowerImpulse, &j->upperImpulse);
  bodyBatch(wg)->hostDirty = true;
}
float f2d_revolute_joint_get_lower_limit(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  return j ? j->lowerAngle : 0.0f;
}
float f2d_revolute_joint_get_upper_limit(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  return j ? j->upperAngle : 0.0f;
}
void f2d_revolute_joint_set_limits(int32_t index1, uint32_t wg, float lower, float upper) {
  Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  if (!j) return;
  if (lower != j->lowerAngle || upper != j->upperAngle) {
    j->lowerAngle = minf(lower, upper);
    j->upperAngle = maxf(lower, upper);
    j->lowerImpulse = 0.0f;
    j->upperImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
int f2d_revolute_joint_is_motor_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  return j ? j->enableMotor : 0;
}
void f2d_revolute_joint_enable_motor(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kRevoluteJoint);
  if (!j) return;
  enableWithReset(j, j->enableMotor, v, &j->motorImpulse);
  bodyBatch(wg)->hostDirty = true;
}
#undef F2D_JOINT_FLOAT
float f2d_revolute_joint_get_motor_torque(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint, &hw);
  return j ? hw->inv_h * j->motorImpulse : 0.0f;
}
// S/revolute_joint.c:72-82
float f2d_revolute_joint_get_angle(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = jointOfType(index1, wg, kRevoluteJoint, &hw);
  if (!j) return 0.0f;
  const float angle = relativeAngle(hw->sims[j->bodyB].xf.q, hw->sims[j->bodyA].xf.q) - j->referenceAngle;
  return unwindAngle(angle);
}

// ---- distance joint (S/distance_joint.c:19-183)
float f2d_distance_joint_get_length(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->targetAngle : 0.0f;
}
void f2d_distance_joint_set_length(int32_t index1, uint32_t wg, float v) {
  Joint* j = jointOfType(index1, wg, kDistanceJoint);
  if (!j) return;
  j->targetAngle = clampf(v, kLinearSlop, kHuge);
  j->linearImpulse.x = 0.0f;
  j->lowerImpulse = 0.0f;
  j->upperImpulse = 0.0f;
  bodyBatch(wg)->hostDirty = true;
}
float f2d_distance_joint_get_current_length(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = jointOfType(index1, wg, kDistanceJoint, &hw);
  if (!j) return 0.0f;
  const Vec2 pA = transformPoint(hw->sims[j->bodyA].xf, j->localOriginAnchorA);
  const Vec2 pB = transformPoint(hw->sims[j->bodyB].xf, j->localOriginAnchorB);
  return length(sub(pB, pA));
}
int f2d_distance_joint_is_spring_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->enableSpring : 0;
}
void f2d_distance_joint_enable_spring(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kDistanceJoint);
  if (!j) return;
  j->enableSpring = v != 0;
  bodyBatch(wg)->hostDirty = true;
}
int f2d_distance_joint_is_limit_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->enableLimit : 0;
}
void f2d_distance_joint_enable_limit(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kDistanceJoint);
  if (!j) return;
  j->enableLimit = v != 0;
  bodyBatch(wg)->hostDirty = true;
}
float f2d_distance_joint_get_min_length(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->lowerAngle : 0.0f;
}
float f2d_distance_joint_get_max_length(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->upperAngle : 0.0f;
}
void f2d_distance_joint_set_length_range(int32_t index1, uint32_t wg, float minLength, float maxLength) {
  Joint* j = jointOfType(index1, wg, kDistanceJoint);
  if (!j) return;
  minLength = clampf(minLength, kLinearSlop, kHuge);
  maxLength = clampf(maxLength, kLinearSlop, kHuge);
  j->lowerAngle = minf(minLength, maxLength);
  j->upperAngle = maxf(minLength, maxLength);
  j->linearImpulse.x = 0.0f;
  j->lowerImpulse = 0.0f;
  j->upperImpulse = 0.0f;
  bodyBatch(wg)->hostDirty = true;
}
int f2d_distance_joint_is_motor_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kDistanceJoint);
  return j ? j->enableMotor : 0;
}
void f2d_distance_joint_enable_motor(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kDistanceJoint);
  if (!j) return;
  if ((v != 0) != (j->enableMotor != 0)) {
    j->enableMotor = v != 0;
    j->motorImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
float f2d_distance_joint_get_motor_force(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = jointOfType(index1, wg, kDistanceJoint, &hw);
  return j ? hw->inv_h * j->motorImpulse : 0.0f;
}

int f2d_wheel_joint_is_spring_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kWheelJoint);
  return j ? j->enableSpring : 0;
}
void f2d_wheel_joint_enable_spring(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kWheelJoint);
  if (!j) return;
  if ((v != 0) != (j->enableSpring != 0)) {
    j->enableSpring = v != 0;
    j->springImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
int f2d_wheel_joint_is_limit_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kWheelJoint);
  return j ? j->enableLimit : 0;
}
void f2d_wheel_joint_enable_limit(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kWheelJoint);
  if (!j) return;
  if ((v != 0) != (j->enableLimit != 0)) {
    j->enableLimit = v != 0;
    j->lowerImpulse = 0.0f;
    j->upperImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
float f2d_wheel_joint_get_lower_limit(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kWheelJoint);
  return j ? j->lowerAngle : 0.0f;
}
float f2d_wheel_joint_get_upper_limit(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kWheelJoint);
  return j ? j->upperAngle : 0.0f;
}
void f2d_wheel_joint_set_limits(int32_t index1, uint32_t wg, float lower, float upper) {
  Joint* j = jointOfType(index1, wg, kWheelJoint);
  if (!j) return;
  if (lower != j->lowerAngle || upper != j->upperAngle) {
    j->lowerAngle = minf(lower, upper);
    j->upperAngle = maxf(lower, upper);
    j->lowerImpulse = 0.0f;
    j->upperImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
int f2d_wheel_joint_is_motor_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = jointOfType(index1, wg, kWheelJoint);
  return j ? j->enableMotor : 0;
}
void f2d_wheel_joint_enable_motor(int32_t index1, uint32_t wg, int v) {
  Joint* j = jointOfType(index1, wg, kWheelJoint);
  if (!j) return;
  if ((v != 0) != (j->enableMotor != 0)) {
    j->enableMotor = v != 0;
    j->motorImpulse = 0.0f;
  }
  bodyBatch(wg)->hostDirty = true;
}
float f2d_wheel_joint_get_motor_torque(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = jointOfType(index1, wg, kWheelJoint, &hw);
  return j ? hw->inv_h * j->motorImpulse : 0.0f;
}

// ---- prismatic joint parameters (f2d_shim.c:1054-1094 -> S/prismatic_joint.c:12-209).
// A call on a joint of another type is ignored (getters return 0).
static Joint* prismaticOf(int32_t index1, uint32_t wg, HostWorld** world = nullptr) {
  int id;
  HostWorld* hw = jointWorld(index1, wg, &id);
  if (!hw || hw->joints[id].type != kPrismaticJoint) return nullptr;
  if (world) *world = hw;
  return &hw->joints[id];
}
static void prismaticChanged(uint32_t wg) {
  Batch* b = bodyBatch(wg);
  if (b) b->hostDirty = true;
}

int f2d_prismatic_joint_is_spring_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->enableSpring : 0;
}
void f2d_prismatic_joint_enable_spring(int32_t index1, uint32_t wg, int v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  if ((v != 0) != (j->enableSpring != 0)) {
    j->enableSpring = v != 0;
    j->springImpulse = 0.0f;
  }
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_spring_hertz(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->hertz : 0.0f;
}
void f2d_prismatic_joint_set_spring_hertz(int32_t index1, uint32_t wg, float v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  j->hertz = v;
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_spring_damping_ratio(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->dampingRatio : 0.0f;
}
void f2d_prismatic_joint_set_spring_damping_ratio(int32_t index1, uint32_t wg, float v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  j->dampingRatio = v;
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_target_translation(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->targetAngle : 0.0f;
}
void f2d_prismatic_joint_set_target_translation(int32_t index1, uint32_t wg, float v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  j->targetAngle = v;
  prismaticChanged(wg);
}
int f2d_prismatic_joint_is_limit_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->enableLimit : 0;
}
void f2d_prismatic_joint_enable_limit(int32_t index1, uint32_t wg, int v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  if ((v != 0) != (j->enableLimit != 0)) {
    j->enableLimit = v != 0;
    j->lowerImpulse = 0.0f;
    j->upperImpulse = 0.0f;
  }
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_lower_limit(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->lowerAngle : 0.0f;
}
float f2d_prismatic_joint_get_upper_limit(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->upperAngle : 0.0f;
}
void f2d_prismatic_joint_set_limits(int32_t index1, uint32_t wg, float lower, float upper) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  if (lower != j->lowerAngle || upper != j->upperAngle) {
    j->lowerAngle = minf(lower, upper);
    j->upperAngle = maxf(lower, upper);
    j->lowerImpulse = 0.0f;
    j->upperImpulse = 0.0f;
  }
  prismaticChanged(wg);
}
int f2d_prismatic_joint_is_motor_enabled(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->enableMotor : 0;
}
void f2d_prismatic_joint_enable_motor(int32_t index1, uint32_t wg, int v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  if ((v != 0) != (j->enableMotor != 0)) {
    j->enableMotor = v != 0;
    j->motorImpulse = 0.0f;
  }
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_motor_speed(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->motorSpeed : 0.0f;
}
void f2d_prismatic_joint_set_motor_speed(int32_t index1, uint32_t wg, float v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  j->motorSpeed = v;
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_max_motor_force(int32_t index1, uint32_t wg) {
  const Joint* j = prismaticOf(index1, wg);
  return j ? j->maxMotorTorque : 0.0f;
}
void f2d_prismatic_joint_set_max_motor_force(int32_t index1, uint32_t wg, float v) {
  Joint* j = prismaticOf(index1, wg);
  if (!j) return;
  j->maxMotorTorque = v;
  prismaticChanged(wg);
}
float f2d_prismatic_joint_get_motor_force(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = prismaticOf(index1, wg, &hw);
  return j ? hw->inv_h * j->motorImpulse : 0.0f;
}
// S/prismatic_joint.c:158-172
float f2d_prismatic_joint_get_translation(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = prismaticOf(index1, wg, &hw);
  if (!j) return 0.0f;
  const Xf xfA = hw->sims[j->bodyA].xf, xfB = hw->sims[j->bodyB].xf;
  const Vec2 axisA = rotate(xfA.q, j->localAxisA);
  const Vec2 pA = transformPoint(xfA, j->localOriginAnchorA);
  const Vec2 pB = transformPoint(xfB, j->localOriginAnchorB);
  return dot(sub(pB, pA), axisA);
}
// S/prismatic_joint.c:174-209: bodies outside the awake set move with zero velocity
float f2d_prismatic_joint_get_speed(int32_t index1, uint32_t wg) {
  HostWorld* hw = nullptr;
  const Joint* j = prismaticOf(index1, wg, &hw);
  if (!j) return 0.0f;
  const BodySim& a = hw->sims[j->bodyA];
  const BodySim& b = hw->sims[j->bodyB];
  const Vec2 axisA = rotate(a.xf.q, j->localAxisA);
  const Vec2 rA = rotate(a.xf.q, sub(j->localOriginAnchorA, a.localCenter));
  const Vec2 rB = rotate(b.xf.q, sub(j->localOriginAnchorB, b.localCenter));
  const Vec2 d = add(sub(b.center, a.center), sub(rB, rA));
  const bool awakeA = a.setIndex == kAwakeSet, awakeB = b.setIndex == kAwakeSet;
  const Vec2 vA = awakeA ? hw->states[j->bodyA].v : Vec2{0.0f, 0.0f};
  const Vec2 vB = awakeB ? hw->states[j->bodyB].v : Vec2{0.0f, 0.0f};
  const float wA = awakeA ? hw->states[j->bodyA].w : 0.0f;
  const float wB = awakeB ? hw->states[j->bodyB].w : 0.0f;
  const Vec2 vRel = sub(add(vB, crossSV(wB, rB)), add(vA, crossSV(wA, rA)));
  return dot(d, crossSV(wA, axisA)) + dot(axisA, vRel);
}

float f2d_body_get_mass(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->bodyMass[id] : 0.0f;
}

float f2d_body_get_rotational_inertia(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->bodyInertia[id] : 0.0f;
}

int f2d_body_get_type(int32_t index1, uint32_t wg) {
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  return hw ? hw->sims[id].type : -1;
}

// f2d_shim.c:412-431
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

static void createShapeFlat(int32_t index1, uint32_t wg, const f2dShapeDef* def, const ShapeGeom& g, int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  int id;
  HostWorld* hw = bodyWorld(index1, wg, &id);
  if (!hw) return;
  Batch* b = bodyBatch(wg);
  if (b->committed) freeDevice(*b);
  int flags = 0;
  if (def->enable_contact_events) flags |= SF_CONTACT_EVENTS;
  if (def->enable_hit_events) flags |= SF_HIT_EVENTS;
  if (def->enable_pre_solve_events) flags |= SF_PRESOLVE;
  if (def->is_sensor) flags |= SF_SENSOR;
  uint64_t cat = (uint64_t)def->category_lo | ((uint64_t)def->category_hi << 32);
  uint64_t mask = (uint64_t)def->mask_lo | ((uint64_t)def->mask_hi << 32);
  int sid = hw->createShape(id, g, def->density, def->friction, def->restitution, def->rolling_resistance,
                            def->tangent_speed, cat, mask, def->group_index, flags, def->invoke_contact_creation != 0,
                            def->update_body_mass != 0);
  b->hostDirty = true;
  out[0] = sid + 1;
  out[1] = (int32_t)((wg & 0xFFFF) | (1u << 16));
}

// f2d_shim.c:459-466
void f2d_create_circle_shape(int32_t index1, uint32_t wg, const f2dShapeDef* def, float cx, float cy, float radius,
                             int32_t* out) {
  ShapeGeom g = {};
  g.type = kCircle;
  g.v[0] = {cx, cy};
  g.radius = radius;
  createShapeFlat(index1, wg, def, g, out);
}

// f2d_shim.c:468-476, b2CreateCapsuleShape S/shape.c:192-201
void f2d_create_capsule_shape(int32_t index1, uint32_t wg, const f2dShapeDef* def, float c1x, float c1y, float c2x,
                              float c2y, float radius, int32_t* out) {
  ShapeGeom g = {};
  Vec2 c1 = {c1x, c1y}, c2 = {c2x, c2y};
  if (distanceSq(c1, c2) <= kLinearSlop * kLinearSlop) {
    g.type = kCircle;
    g.v[0] = lerp(c1, c2, 0.5f);
  } else {
    g.type = kCapsule;
    g.v[0] = c1;
    g.v[1] = c2;
  }
  g.radius = radius;
  createShapeFlat(index1, wg, def, g, out);
}

// f2d_shim.c:478-486
void f2d_create_segment_shape(int32_t index1, uint32_t wg, const f2dShapeDef* def, float p1x, float p1y, float p2x,
                              float p2y, int32_t* out) {
  ShapeGeom g = {};
  g.type = kSegment;
  g.v[0] = {p1x, p1y};
  g.v[1] = {p2x, p2y};
  createShapeFlat(index1, wg, def, g, out);
}

// f2d_shim.c:488-500
void f2d_create_box_shape(int32_t index1, uint32_t wg, const f2dShapeDef* def, float half_width, float half_height,
                          float cx, float cy, float qc, float qs, float radius, int32_t* out) {
  ShapeGeom g = makeOffsetRoundedBox(half_width, half_height, {cx, cy}, {qc, qs}, radius);
  createShapeFlat(index1, wg, def, g, out);
}

// f2d_shim.c:502-515: returns 0 when the hull is degenerate
int f2d_create_polygon_shape(int32_t index1, uint32_t wg, const f2dShapeDef* def, const float* points, int count,
                             float radius, int32_t* out) {
  Hull hull = computeHull(reinterpret_cast<const Vec2*>(points), count);
  if (hull.count == 0) {
    out[0] = 0;
    out[1] = 0;
    return 0;
  }
  ShapeGeom g = makePolygon(hull, radius);
  createShapeFlat(index1, wg, def, g, out);
  return 1;
}

// A joint with collideConnected off destroys the contacts between its bodies when it is
// created (b2DestroyContactsBetweenBodies, S/joint.c:313-345; not for mouse and filter
// joints). That edit of a live world's contact, island and colour state is not restated
// on the host, so such a creation is refused while the bodies share a contact.
static bool refuseContactsBetween(const HostWorld& hw, int a, int b, int collideConnected) {
  if (collideConnected) return false;
  int key = hw.sims[a].headContactKey;
  while (key >= 0) {
    const Contact& c = hw.contacts[key >> 1];
    const int other = (key & 1) == 0 ? c.bodyB : c.bodyA;
    if (other == b) {
      fprintf(stderr, "forge2d_amd: a joint without collideConnected between bodies %d and %d, which share "
                      "contact %d, is not supported (the reference destroys that contact)\n", a, b, key >> 1);
      return true;
    }
    key = (key & 1) == 0 ? c.nextA : c.nextB;
  }
  return false;
}

// f2d_shim.c:816-843 (revolute subset of the def)
void f2d_create_revolute_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                               float lax, float lay, float lbx, float lby, float reference_angle, float target_angle,
                               int enable_spring, float hertz, float damping_ratio, int enable_limit,
                               float lower_angle, float upper_angle, int enable_motor, float max_motor_torque,
                               float motor_speed, float draw_size, int collide_connected, int32_t* out) {
  (void)draw_size;
  (void)b_wg;
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createRevolute(ia, ib, {lax, lay}, {lbx, lby}, reference_angle, target_angle, enable_spring != 0, hertz,
                              damping_ratio, enable_limit != 0, lower_angle, upper_angle, enable_motor != 0,
                              max_motor_torque, motor_speed, collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:787-814
void f2d_create_prismatic_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                                float lax, float lay, float lbx, float lby, float axis_x, float axis_y,
                                float reference_angle, float target_translation, int enable_spring, float hertz,
                                float damping_ratio, int enable_limit, float lower_translation,
                                float upper_translation, int enable_motor, float max_motor_force, float motor_speed,
                                int collide_connected, int32_t* out) {
  (void)b_wg;
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createPrismatic(ia, ib, {lax, lay}, {lbx, lby}, {axis_x, axis_y}, reference_angle, target_translation,
                               enable_spring != 0, hertz, damping_ratio, enable_limit != 0, lower_translation,
                               upper_translation, enable_motor != 0, max_motor_force, motor_speed,
                               collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:716-740
void f2d_create_distance_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                               float lax, float lay, float lbx, float lby, float length, int enable_spring, float hertz,
                               float damping_ratio, int enable_limit, float min_length, float max_length,
                               int enable_motor, float max_motor_force, float motor_speed, int collide_connected,
                               int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createDistance(ia, ib, {lax, lay}, {lbx, lby}, length, enable_spring != 0, hertz, damping_ratio,
                              enable_limit != 0, min_length, max_length, enable_motor != 0, max_motor_force,
                              motor_speed, collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:751-768
void f2d_create_motor_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                            float offset_x, float offset_y, float angular_offset, float max_force, float max_torque,
                            float correction_factor, int collide_connected, int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createMotor(ia, ib, {offset_x, offset_y}, angular_offset, max_force, max_torque, correction_factor,
                           collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:770-785
void f2d_create_mouse_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                            float target_x, float target_y, float hertz, float damping_ratio, float max_force,
                            int collide_connected, int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createMouse(ia, ib, {target_x, target_y}, hertz, damping_ratio, max_force, collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:742-749
void f2d_create_filter_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg,
                             int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createFilter(ia, ib);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:844-862
void f2d_create_weld_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg, float lax,
                           float lay, float lbx, float lby, float reference_angle, float linear_hertz,
                           float angular_hertz, float linear_damping_ratio, float angular_damping_ratio,
                           int collide_connected, int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createWeld(ia, ib, {lax, lay}, {lbx, lby}, reference_angle, linear_hertz, angular_hertz,
                          linear_damping_ratio, angular_damping_ratio, collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// f2d_shim.c:864-888
void f2d_create_wheel_joint(uint32_t w, int32_t a_index1, uint32_t a_wg, int32_t b_index1, uint32_t b_wg, float lax,
                            float lay, float lbx, float lby, float axis_x, float axis_y, int enable_spring, float hertz,
                            float damping_ratio, int enable_limit, float lower_translation, float upper_translation,
                            int enable_motor, float max_motor_torque, float motor_speed, int collide_connected,
                            int32_t* out) {
  out[0] = 0;
  out[1] = 0;
  Batch* b = batchOf(w);
  int ia, ib;
  HostWorld* hw = bodyWorld(a_index1, a_wg, &ia);
  if (!b || !hw || !bodyWorld(b_index1, b_wg, &ib)) return;
  if (refuseContactsBetween(*hw, ia, ib, collide_connected)) return;
  if (b->committed) freeDevice(*b);
  int id = hw->createWheel(ia, ib, {lax, lay}, {lbx, lby}, {axis_x, axis_y}, enable_spring != 0, hertz, damping_ratio,
                           enable_limit != 0, lower_translation, upper_translation, enable_motor != 0,
                           max_motor_torque, motor_speed, collide_connected != 0);
  b->hostDirty = true;
  out[0] = id + 1;
  out[1] = (int32_t)(((w & 0xFFFF) - 1) | ((uint32_t)hw->jointGen[id] << 16));
}

// Batch handle behind a single world, for the batch ABI (export, counters).
void* f2d_world_batch(uint32_t w) { return batchOf(w); }

}  // extern "C"

// ---- events through the flat per-world ABI (F/native/wasm/f2d_shim.c:1186-1300)

static uint32_t shapeWgOf(uint32_t w) { return ((w & 0xFFFF) - 1) | (1u << 16); }

extern "C" {

void f2d_world_get_contact_event_counts(uint32_t w, int32_t* out) {
  Batch* b = batchOf(w);
  if (!b) return;
  int c[3];
  contactEventCounts(*b, 0, c);
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
}
int f2d_world_get_contact_begin_events(uint32_t w, double* out) {
  Batch* b = batchOf(w);
  return b ? beginEvents(*b, 0, out, shapeWgOf(w)) : 0;
}
int f2d_world_get_contact_end_events(uint32_t w, double* out) {
  Batch* b = batchOf(w);
  return b ? endEvents(*b, 0, out, shapeWgOf(w)) : 0;
}
int f2d_world_get_contact_hit_events(uint32_t w, double* out) {
  Batch* b = batchOf(w);
  return b ? hitEvents(*b, 0, out, shapeWgOf(w)) : 0;
}
int f2d_world_get_body_event_count(uint32_t w) {
  Batch* b = batchOf(w);
  return b ? bodyEvents(*b, 0, nullptr, (int)(w & 0xFFFF) - 1) : 0;
}
int f2d_world_get_body_events(uint32_t w, double* out) {
  Batch* b = batchOf(w);
  return b ? bodyEvents(*b, 0, out, (int)(w & 0xFFFF) - 1) : 0;
}

}  // extern "C"

// ---- world queries (S/world.c:2050-2308, 2640-2747) ------------------------------

// the public header's query records and host callbacks (include/forge2d_amd.h)
typedef float (*f2dHostCastRayFn)(int32_t shape_index1, uint32_t shape_wg, float px, float py, float nx, float ny,
                                  float fraction);
typedef int (*f2dHostOverlapFn)(int32_t shape_index1, uint32_t shape_wg);
typedef struct f2dQuery {
  int32_t world;
  int32_t kind;
  uint32_t category_lo, category_hi, mask_lo, mask_hi;
  float p[6];
} f2dQuery;
typedef struct f2dQueryHit {
  int32_t shape;
  float px, py, nx, ny, fraction;
} f2dQueryHit;

namespace f2d {

// The query grouping scratch for n queries: the order, then a count and first slot
// per world (and one for queries naming no world of the batch).
static int ensureGroupBuffer(Batch& b, long n) {
  const long groupInts = n + (long)b.worlds.size() + 1;
  if (n > 0 && groupInts > b.dGroupCap) {
    if (b.dGroup) (void)hipFree(b.dGroup);
    b.dGroup = nullptr;
    b.dGroupCap = 0;
    const long c = std::max<long>(groupInts, 1024);
    F2D_HIP(hipMalloc(&b.dGroup, sizeof(int) * c));
    b.dGroupCap = c;
  }
  return 0;
}

static int ensureQueryBuffers(Batch& b, long n, long hitCount) {
  if (ensureGroupBuffer(b, n) != 0) return -1;
  if (n > b.dQueryCap) {
    if (b.dQueries) (void)hipFree(b.dQueries);
    if (b.dCounts) (void)hipFree(b.dCounts);
    b.dQueries = nullptr;
    b.dCounts = nullptr;
    b.dQueryCap = 0;
    const long c = std::max<long>(n, 64);
    F2D_HIP(hipMalloc(&b.dQueries, sizeof(Query) * c));
    F2D_HIP(hipMalloc(&b.dCounts, sizeof(int) * c));
    b.dQueryCap = c;
  }
  if (hitCount > b.dHitCap) {
    if (b.dHits) (void)hipFree(b.dHits);
    b.dHits = nullptr;
    b.dHitCap = 0;
    const long c = std::max<long>(hitCount, 256);
    F2D_HIP(hipMalloc(&b.dHits, sizeof(QueryHit) * c));
    b.dHitCap = c;
  }
  return 0;
}

// Rays and overlaps in one launch; explosions afterwards in rounds: the k-th explosion
// of a world runs in round k, so explosions of one world apply in query order and
// those of different worlds side by side.
static int runQueries(Batch& b, const Query* q, int n, QueryHit* hits, int* counts, int cap) {
  if (n < 0 || cap < 1) return -1;
  if (n == 0) return 0;
  if (b.hostDirty || !b.committed) {
    if (upload(b) != 0) return -1;
  }
  std::vector<int> roundOf(b.worlds.size(), 0);
  std::vector<std::vector<Query>> rounds;
  for (int i = 0; i < n; ++i) {
    if (q[i].kind != QK_EXPLODE || q[i].world < 0 || q[i].world >= (int)b.worlds.size()) continue;
    const int r = roundOf[q[i].world]++;
    if ((int)rounds.size() <= r) rounds.resize(r + 1);
    rounds[r].push_back(q[i]);
  }
  long explosions = 0;
  for (const auto& r : rounds) explosions += (long)r.size();
  if (ensureQueryBuffers(b, n + explosions, (long)n * cap) != 0) return -1;
  F2D_HIP(hipMemcpyAsync(b.dQueries, q, sizeof(Query) * n, hipMemcpyHostToDevice, b.stream));
  if (f2d_launch_queries(&b.dev.view, b.dGroup, b.dQueries, n, b.dHits, b.dCounts, cap, b.stream) != 0) return -1;
  long off = n;
  for (const auto& r : rounds) {
    F2D_HIP(hipMemcpyAsync(b.dQueries + off, r.data(), sizeof(Query) * r.size(), hipMemcpyHostToDevice, b.stream));
    if (f2d_launch_explosions(&b.dev.view, b.dQueries + off, (int)r.size(), b.stream) != 0) return -1;
    off += (long)r.size();
  }
  F2D_HIP(hipMemcpyAsync(counts, b.dCounts, sizeof(int) * n, hipMemcpyDeviceToHost, b.stream));
  F2D_HIP(hipStreamSynchronize(b.stream));
  // copy back only the hit slots that hold results
  long used = 0;
  for (int i = 0; i < n; ++i) used = std::max<long>(used, (long)i * cap + std::min(std::max(counts[i], 0), cap));
  if (used > 0) F2D_HIP(hipMemcpy(hits, b.dHits, sizeof(QueryHit) * used, hipMemcpyDeviceToHost));
  b.deviceNewer = true;  // pending tree enlargements applied, explosions
  return 0;
}

static f2dHostCastRayFn g_castRayFn = nullptr;
static f2dHostOverlapFn g_overlapFn = nullptr;

static int flatQuery(Batch& b, const Query& q, std::vector<QueryHit>& out) {
  const int cap = std::max((int)b.worlds[0].shapes.size(), 1);
  out.resize(cap);
  int count = 0;
  if (runQueries(b, &q, 1, out.data(), &count, cap) != 0) return -1;
  out.resize(std::max(count, 0));
  return 0;
}

// the tree (static 0, kinematic 1, dynamic 2) a shape's proxy lives in
static int treeOf(const HostWorld& hw, int shape) { return hw.shapes[shape].proxyKey & 3; }

}  // namespace f2d

extern "C" {

int f2dg_query(void* h, const f2dQuery* queries, int n, f2dQueryHit* hits, int32_t* counts, int cap) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || (n > 0 && (!queries || !hits || !counts))) return -1;
  static_assert(sizeof(f2dQuery) == sizeof(Query) && sizeof(f2dQueryHit) == sizeof(QueryHit), "query records");
  return runQueries(*b, reinterpret_cast<const Query*>(queries), n, reinterpret_cast<QueryHit*>(hits), counts, cap);
}

int f2dg_query_device(void* h, const void* queries, int n, void* hits, void* counts, int cap) {
  Batch* b = static_cast<Batch*>(h);
  if (!b || n < 0 || cap < 1) return -1;
  if (n == 0) return 0;
  if (b->hostDirty || !b->committed) {
    if (upload(*b) != 0) return -1;
  }
  if (ensureGroupBuffer(*b, n) != 0) return -1;
  b->deviceNewer = true;  // pending tree enlargements applied
  return f2d_launch_queries(&b->dev.view, b->dGroup, static_cast<const Query*>(queries), n,
                            static_cast<QueryHit*>(hits), static_cast<int*>(counts), cap, b->stream);
}

void f2d_set_query_callbacks(f2dHostCastRayFn cast_ray, f2dHostOverlapFn overlap) {
  g_castRayFn = cast_ray;
  g_overlapFn = overlap;
}

static Query flatQueryRecord(int kind, float a, float b_, float c, float d, uint32_t catLo, uint32_t catHi,
                             uint32_t maskLo, uint32_t maskHi) {
  Query q;
  std::memset(&q, 0, sizeof(q));
  q.world = 0;
  q.kind = kind;
  q.catLo = catLo;
  q.catHi = catHi;
  q.maskLo = maskLo;
  q.maskHi = maskHi;
  q.p[0] = a;
  q.p[1] = b_;
  q.p[2] = c;
  q.p[3] = d;
  return q;
}

// f2d_shim.c:1313-1333 -> b2World_CastRayClosest
int f2d_world_cast_ray_closest(uint32_t w, float ox, float oy, float tx, float ty, uint32_t category_lo,
                               uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi, double* out) {
  Batch* b = batchOf(w);
  if (!b) return 0;
  Query q = flatQueryRecord(QK_RAY_CLOSEST, ox, oy, tx, ty, category_lo, category_hi, mask_lo, mask_hi);
  QueryHit hit;
  int count = 0;
  if (runQueries(*b, &q, 1, &hit, &count, 1) != 0 || count < 1) return 0;
  out[0] = hit.shape + 1;
  out[1] = shapeWgOf(w);
  out[2] = hit.px;
  out[3] = hit.py;
  out[4] = hit.nx;
  out[5] = hit.ny;
  out[6] = hit.fraction;
  return 1;
}

// f2d_shim.c:1343-1350 -> b2World_CastRay (S/world.c:2222-2254): the hits reach the
// host callback in the trees' walk order; a return of 0 ends the cast, a value in
// [0, 1] becomes the world's fraction (the next tree's ray length), a value in
// (0, current] clips the current tree's ray so hits beyond it are skipped
void f2d_world_cast_ray(uint32_t w, float ox, float oy, float tx, float ty, uint32_t category_lo,
                        uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi) {
  Batch* b = batchOf(w);
  if (!b || !g_castRayFn) return;
  Query q = flatQueryRecord(QK_RAY_ALL, ox, oy, tx, ty, category_lo, category_hi, mask_lo, mask_hi);
  std::vector<QueryHit> hits;
  if (flatQuery(*b, q, hits) != 0) return;
  const HostWorld& hw = b->worlds[0];
  // the hits come in the trees' walk order with the whole ray (maxFraction 1): the
  // reference's walk with the callback's fractions visits the same hits in the same
  // order, minus those beyond the fraction the callback last clipped the ray to
  float worldFraction = 1.0f;
  for (int tree = kStatic; tree <= kDynamic; ++tree) {
    float maxFraction = worldFraction;
    for (const QueryHit& hit : hits) {
      if (treeOf(hw, hit.shape) != tree || hit.fraction > maxFraction) continue;
      const float value = g_castRayFn(hit.shape + 1, shapeWgOf(w), hit.px, hit.py, hit.nx, hit.ny, hit.fraction);
      if (0.0f <= value && value <= 1.0f) worldFraction = value;
      if (value == 0.0f) return;
      if (0.0f < value && value <= maxFraction) maxFraction = value;
    }
    if (worldFraction == 0.0f) return;
  }
}

// f2d_shim.c:1357-1367 -> b2World_OverlapAABB (S/world.c:2071-2105): per tree in its
// depth-first order; a false return ends that tree's walk
void f2d_world_overlap_aabb(uint32_t w, float lx, float ly, float ux, float uy, uint32_t category_lo,
                            uint32_t category_hi, uint32_t mask_lo, uint32_t mask_hi) {
  Batch* b = batchOf(w);
  if (!b || !g_overlapFn) return;
  Query q = flatQueryRecord(QK_OVERLAP_AABB, lx, ly, ux, uy, category_lo, category_hi, mask_lo, mask_hi);
  std::vector<QueryHit> hits;
  if (flatQuery(*b, q, hits) != 0) return;
  const HostWorld& hw = b->worlds[0];
  for (int tree = kStatic; tree <= kDynamic; ++tree) {
    for (const QueryHit& hit : hits) {
      if (treeOf(hw, hit.shape) != tree) continue;
      if (!g_overlapFn(hit.shape + 1, shapeWgOf(w))) break;
    }
  }
}

// f2d_shim.c:177-189 -> b2World_Explode
void f2d_world_explode(uint32_t w, uint32_t mask_lo, uint32_t mask_hi, float x, float y, float radius,
                       float falloff, float impulse_per_length) {
  Batch* b = batchOf(w);
  if (!b) return;
  Query q = flatQueryRecord(QK_EXPLODE, x, y, radius, falloff, 0u, 0u, mask_lo, mask_hi);
  q.p[4] = impulse_per_length;
  QueryHit unused;
  int count = 0;
  (void)runQueries(*b, &q, 1, &unused, &count, 1);
}

}  // extern "C"
