// Device data layout of a batch of worlds.
//
// All worlds of a batch live in shared flat arrays; a WorldHeader holds each
// world's offsets into them. Bodies and shapes are array-of-structs (the solver
// gathers whole body states by index; per-body passes read a whole record).
// Contact constraints built each step are struct-of-arrays (ConstraintSoA), so
// the lanes of a colour stage read each field with one coalesced access.
//
// Ids are world-local: body id, shape id and contact id index the world's slice
// of the flat arrays, exactly as bodyId/shapeId/contactId index the reference's
// b2World arrays (S/world.h).
#pragma once

#include "b2math.hpp"
#include "tree.hpp"

namespace f2d {

enum BodyType : int { kStatic = 0, kKinematic = 1, kDynamic = 2 };
enum ShapeType : int { kCircle = 0, kCapsule = 1, kSegment = 2, kPolygon = 3 };

// body flags
enum : int {
  BF_BULLET = 1,
  BF_FIXED_ROTATION = 2,
  BF_ALLOW_FAST_ROTATION = 4,
  BF_ENABLE_SLEEP = 8,
  BF_IS_FAST = 16,
  BF_SPEED_CAPPED = 32,
  BF_ENLARGED = 64,
  BF_CAPPED_STEP = 128,  // speed capped during the current step (reported at finalize)
  BF_FELL_ASLEEP = 256,  // put to sleep by the last step (its move event's fellAsleep)
};
// shape flags
enum : int {
  SF_CONTACT_EVENTS = 1,
  SF_HIT_EVENTS = 2,
  SF_ENLARGED = 4,   // fat AABB grew in the last finalize (enlargedAABB)
  SF_MOVED = 8,      // proxy is in the move buffer for the next pair update
  SF_PRESOLVE = 16,
  SF_SENSOR = 32,
  SF_TREE_PENDING = 64,  // the step enlarged this proxy's leaf; the tree kernel applies it
};
// contact flags (b2ContactFlags, S/contact.h:17-40) and sim flags (:66-86)
enum : int {
  CF_TOUCHING = 0x1,
  CF_CONTACT_EVENTS = 0x4,  // b2_contactEnableContactEvents (0x2 is the unused hit-event flag)
  CF_OFF_AWAKE = 0x100,     // not in the awake set (disabled or sleeping set): not collided
  CF_REF_MASK = 0x7,        // the bits that are the reference's b2ContactFlags
  SIM_TOUCHING = 0x00010000,
  SIM_DISJOINT = 0x00020000,
  SIM_STARTED = 0x00040000,
  SIM_STOPPED = 0x00080000,
  SIM_HIT_EVENT = 0x00100000,
  SIM_PRESOLVE = 0x00200000,
};
// world flags
enum : int { WF_SLEEP = 1, WF_CONTINUOUS = 2, WF_WARM = 4 };
// error bits reported back to the host
enum : int {
  ERR_CONTACT_CAP = 1,
  ERR_PAIR_CAP = 2,
  ERR_OVERFLOW_CAP = 4,
  ERR_SCRATCH = 8,
  ERR_TREE = 16,  // a broad-phase tree and its ranks disagreed; the tree was not rebuilt
};

// b2BodyState, S/body.h:100-113 (32 bytes, identical field order)
struct alignas(16) BodyState {
  Vec2 v;
  float w;
  int flags;
  Vec2 dp;
  Rot dq;
};

// b2BodySim (S/body.h:118-159) plus the b2Body fields the step reads.
struct alignas(16) BodySim {
  Xf xf;            // transform of the body origin
  Vec2 center;      // world centre of mass
  Vec2 center0;     // centre of mass at the start of the step (TOI)
  Rot rotation0;
  Vec2 localCenter;
  Vec2 force;
  float torque;
  float invMass, invI;
  float minExtent, maxExtent;
  float linearDamping, angularDamping, gravityScale;
  int type;
  int flags;
  int headShape;       // first shape of the body's shape list (b2Body::headShapeId)
  int headContactKey;  // b2Body::headContactKey
  int contactCount;
  int headJointKey;
  int jointCount;
  float sleepThreshold;
  float sleepTime;
  int islandId, islandPrev, islandNext;  // island graph (b2Body, S/body.h:40-47)
  int setIndex;    // solver set: kStaticSet, kAwakeSet or a sleeping set
  int localIndex;  // index within the set (the awake set's order is the sim order)
};

struct alignas(16) ShapeGeom {
  int type;
  int count;     // polygon vertex count
  float radius;  // circle/capsule/polygon radius
  float pad;
  Vec2 centroid;
  Vec2 v[kMaxPolygonVertices];  // circle: v[0] centre; capsule/segment: v[0], v[1]
  Vec2 n[kMaxPolygonVertices];
};

struct alignas(16) Shape {
  AABB aabb;
  AABB fat;
  int body;
  int next;       // next shape on the same body (b2Shape::nextShapeId)
  int proxyKey;   // (proxyId << 2) | bodyType, S/broad_phase.h:18-21
  int flags;
  float friction, restitution, rollingResistance, tangentSpeed;
  uint32_t catLo, catHi, maskLo, maskHi;
  int group;
  int moveRank;   // order of this proxy in the move buffer when it moves
  int pad[2];
};

// b2ManifoldPoint, I/collision.h:498-536
struct ManifoldPoint {
  Vec2 point, anchorA, anchorB;
  float separation, normalImpulse, tangentImpulse, totalNormalImpulse, normalVelocity;
  int id;  // uint16 id in the low bits, persisted flag in bit 16
};

struct Manifold {
  Vec2 normal;
  float rollingImpulse;
  int pointCount;
  ManifoldPoint points[2];
};

// One slot per contact id: b2Contact + b2ContactSim (S/contact.h:42-131).
struct alignas(16) Contact {
  int shapeA, shapeB;   // -1 when the slot is free
  int flags;            // CF_*
  int simFlags;         // SIM_*
  int color;            // constraint-graph colour or -1
  int localIndex;       // index within the overflow list when color == overflow
  int bodyA, prevA, nextA;  // edge 0 (S/contact.h b2ContactEdge)
  int bodyB, prevB, nextB;  // edge 1
  float friction, restitution, rollingResistance, tangentSpeed;
  Manifold m;
};

// Revolute joint slot (b2Joint + b2JointSim + b2RevoluteJoint, S/joint.h:27-60, 162-190, 247-281)
// b2JointType values of the joints built here (I/types.h)
enum JointType : int {
  kDistanceJoint = 0,
  kFilterJoint = 1,
  kMotorJoint = 2,
  kMouseJoint = 3,
  kPrismaticJoint = 4,
  kRevoluteJoint = 5,
  kWeldJoint = 6,
  kWheelJoint = 7
};

// One joint (b2Joint + b2JointSim, S/joint.h). The joint types share fields
// (S/joint.h b2PrismaticJoint, b2WeldJoint, b2WheelJoint):
//   prismatic: targetAngle = target translation, lowerAngle/upperAngle = translation
//     limits, maxMotorTorque = maximum motor force, linearImpulse = (perpendicular,
//     angular) impulse of its block constraint;
//   weld: hertz/dampingRatio = linear stiffness, springSoftness = linear softness,
//     springImpulse = angular impulse;
//   wheel: lowerAngle/upperAngle = translation limits, linearImpulse.x =
//     perpendicular impulse;
//   distance: targetAngle = rest length, lowerAngle/upperAngle = length range,
//     maxMotorTorque = maximum motor force, springSoftness = distance softness,
//     linearImpulse.x = the length impulse;
//   motor: localAxisA = linear offset, referenceAngle = angular offset,
//     maxMotorTorque = maximum force, angularHertz = maximum torque,
//     angularDampingRatio = correction factor, springImpulse = angular impulse;
//   mouse: localAxisA = target, maxMotorTorque = maximum force, springImpulse =
//     angular impulse;
//   motor and mouse keep their prepared 2x2 linear mass in axisA (first column) and
//     (perpMass, motorMass) (second column).
struct alignas(16) Joint {
  int bodyA, bodyB;
  int color, localIndex;
  int prevA, nextA, prevB, nextB;  // body joint list edges
  int collideConnected;
  int enableSpring, enableMotor, enableLimit;
  Vec2 localOriginAnchorA, localOriginAnchorB;
  float referenceAngle, targetAngle, hertz, dampingRatio, maxMotorTorque, motorSpeed, lowerAngle, upperAngle;
  float constraintHertz, constraintDampingRatio;
  Vec2 linearImpulse;
  float springImpulse, motorImpulse, lowerImpulse, upperImpulse;
  // per-step prepared data
  float invMassA, invMassB, invIA, invIB;
  Softness constraintSoftness, springSoftness;
  Vec2 anchorA, anchorB, deltaCenter;
  float deltaAngle, axialMass;
  int islandId, islandPrev, islandNext;  // island graph (b2Joint, S/joint.h:44-48)
  int setIndex;
  int simA, simB;  // solver state index of each body, -1 outside the awake set (set by prepare)
  int type;        // JointType
  int pad;
  Vec2 localAxisA, axisA;  // prismatic, wheel: unit axis in body A's frame; rotated by prepare
  float angularHertz, angularDampingRatio;  // weld
  float perpMass, motorMass;                // wheel (prepared)
  Softness angularSoftness;                 // weld (prepared)
  int pad2;
};

// Per-world header. Offsets index the batch's flat arrays.
struct alignas(16) WorldHeader {
  int bodyOffset, bodyCount;
  int shapeOffset, shapeCount;
  int contactOffset, contactCap;
  int contactCount;      // slots in use = next fresh id (contact array count)
  int freeCount;         // LIFO free-id stack depth (stack at freeOffset)
  int freeOffset;
  int jointOffset, jointCount;
  int pairOffset, pairCap;
  int colorOffset, colorWords;   // (kOverflowIndex) bitsets of colorWords uint32 each
  int overflowOffset, overflowCount;
  int constraintOffset;          // constraint SoA region, contactCap lanes
  int scratchOffset, scratchCap; // int scratch for large-world sorts
  int flags;
  int error;
  int stepIndex;
  int activeColors;              // colours used in the last solve (statistics)
  Vec2 gravity;
  float restitutionThreshold, maxLinearSpeed, contactHertz, contactDampingRatio, maxContactPushSpeed;
  float contactSpeed;            // world->contactSpeed, set each step
  float inv_h;                   // world->inv_h
  float jointHertz, jointDampingRatio;
  int fastCount;                 // fast bodies listed by finalize for continuous collision
  int staticCount;               // static proxies listed at staticShapes[shapeOffset ...]
  int beginCount, endCount;      // contact begin/end touch events of the last collide
  float hitEventThreshold;       // b2WorldDef::hitEventThreshold
  // islands and solver sets (S/island.c, S/solver_set.c); region at islandOffset of
  // the batch's island array, laid out by islandRegion()
  int islandOffset, islandCap;   // cap: islands a world can hold (its non-static bodies)
  int islandCount;               // island array count = island id-pool next index
  int islandFreeCount;
  int awakeIslandCount;
  int splitIslandId;             // candidate chosen by the last step's finalize
  int setCount, setFreeCount;    // solver-set id pool (next index, free list depth)
  int awakeBodyCount;
  int moveBodyCount;             // awake bodies finalized by the last step (move events)
  int sleepPending;              // finalize ran with sleep enabled: island sleep follows
  int mergePending;              // an island got a union-find parent since the last merge
  int sweepCount;                // shapes the multi-workgroup sweep launch sweeps (0: none)
  int sweepPairs;                // new pairs that launch found
  int colorListOffset;           // colour arrays of colours 0..10 (kColorListHeader counts, then
                                 // one array of bodyCount slots per colour) in the batch's colorLists
  unsigned long long splitKey;   // finalize: sleepiest body of an island with removed constraints
};

// Per-world colour arrays (the b2GraphColor contact arrays, S/constraint_graph.h:22-45):
// a header of per-colour counts, then colour c's contact ids in array order at
// kColorListHeader + c * bodyCount. A colour other than the overflow holds each
// non-static body once, so bodyCount slots suffice.
constexpr int kColorListHeader = 16;
F2D_HD long colorListInts(int bodyCount) { return kColorListHeader + (long)kOverflowIndex * bodyCount; }

// Solver sets, S/solver_set.h
enum : int { kStaticSet = 0, kDisabledSet = 1, kAwakeSet = 2, kFirstSleepingSet = 3 };

// b2Island, S/island.h:27-58 (also the snapshot's island record)
struct Island {
  int id, setIndex, localIndex;
  int headBody, tailBody, bodyCount;
  int headContact, tailContact, contactCount;
  int headJoint, tailJoint, jointCount;
  int parent;          // union-find parent, -1 for a root
  int removeCount;     // constraints removed since the island was built (split candidate)
  int reserved[2];
};

// Per-world island region (ints): island records, the island id pool's free list,
// the awake set's island order, the solver-set id pool's free list, the awake set's
// body order, the body order of the last finalize (move events), the finalize's
// awake-island bits, and the split's depth-first stack and seed list.
struct IslandRegion {
  long islands, islandFree, awakeIslands, setFree, awakeBodies, moveOrder, awakeBits, stack, seeds, total;
};
F2D_HD IslandRegion islandRegion(int cap, int bodies) {
  IslandRegion r;
  r.islands = 0;
  r.islandFree = r.islands + 16L * cap;
  r.awakeIslands = r.islandFree + cap;
  r.setFree = r.awakeIslands + cap;
  r.awakeBodies = r.setFree + cap + 4;
  r.moveOrder = r.awakeBodies + bodies;
  r.awakeBits = r.moveOrder + bodies;
  r.stack = r.awakeBits + (cap + 31) / 32;
  r.seeds = r.stack + bodies;
  r.total = (r.seeds + bodies + 3) & ~3L;
  return r;
}

// Per-step scalars, S/world.c:743-771
struct StepParams {
  float dt, inv_dt, h, inv_h;
  int subSteps;
  int phases;
};
enum : int {
  PH_PAIRS = 1,
  PH_COLLIDE = 2,
  PH_STATE = 4,
  PH_SOLVE = 8,
  PH_FINALIZE = 16,
  PH_ALL = 31,
  PH_NO_COUNT = 32,        // collide launch of a split step: the solve launch counts the step
  PH_CONTINUOUS = 64,      // launch the continuous-collision kernel on its own
  PH_NO_CONTINUOUS = 128,  // finalize without the continuous-collision launch (timing splits)
  PH_PAIRS_PREP = 256,     // pair update of a large world: sort only, the sweep runs in its own launch
  PH_PAIRS_FINISH = 512,   // pair update of a large world: create the pairs the sweep launch found
  PH_NO_TREE = 1024,       // collide without the tree launch (timing splits)
  PH_TREE = 2048,          // the tree launch on its own (timing splits)
  PH_CONTACTS_DONE = 4096, // collide launch of a large world: the narrow-phase launch updated the contacts
  PH_SOLVE_PREP = 8192,    // solve launch of a large world: islands and colour buckets only, the stages
                           // run as launches over many workgroups
};

// A world query of the batch query ABI (f2dQuery in the public header). Rays:
// p = origin x/y, translation x/y; AABB overlaps: p = lower x/y, upper x/y;
// explosions: p = position x/y, radius, falloff, impulse per length, with the mask
// in maskLo/maskHi.
enum QueryKind : int { QK_RAY_CLOSEST = 1, QK_RAY_ALL = 2, QK_OVERLAP_AABB = 3, QK_EXPLODE = 4 };
struct Query {
  int world;
  int kind;
  uint32_t catLo, catHi, maskLo, maskHi;
  float p[6];
};
// A ray hit (shape id, point, normal, fraction) or an overlapping shape (shape id)
struct QueryHit {
  int shape;
  float px, py, nx, ny, fraction;
};
static_assert(sizeof(Query) == 48 && sizeof(QueryHit) == 24, "query records");

// Launch shape of a batch: workgroup size and dynamic LDS bytes per kernel.
// HBM work region of tree t of a world whose node region starts at `offset`
F2D_HD long treeWorkOffset(int offset, int world, int t) { return 25L * offset + 160L * (3L * world + t); }

struct LaunchConfig {
  int block;       // 64, 128 or 256 lanes per world
  int collideLds;  // sort-and-sweep + pair-sort LDS
  int solveLds;    // body-state LDS
  int continuous;  // some world of the batch has continuous collision enabled
  int sweepChunks; // > 0: worlds whose sort arrays exceed LDS sweep in their own launch, this
                   // many 256-shape chunks per world, and update their contacts in a
                   // narrow-phase launch over many workgroups
  int treeLds;     // tree rebuild work arrays (worlds whose trees need more use HBM)
  int treeBlock;   // lanes of the tree launch (64, 128, 256 or 1024)
  int continuousChunks = 0;  // > 0: the non-bullet sweeps of each world spread over this many
                             // workgroups, then bullets and the sleep phase in a second launch
  int wideChunks = 0;        // > 0: the solve stages of each world as launches of this many
                             // 256-lane workgroups
};

// Contact constraint lanes (one constraint per lane), S/contact_solver.c:1034-1064.
enum ConstraintField : int {
  CN_INDEX_A = 0, CN_INDEX_B, CN_SLOT,
  CN_INV_MASS_A, CN_INV_MASS_B, CN_INV_I_A, CN_INV_I_B,
  CN_NX, CN_NY, CN_FRICTION, CN_TANGENT_SPEED, CN_ROLLING_RES, CN_ROLLING_MASS, CN_ROLLING_IMPULSE,
  CN_BIAS_RATE, CN_MASS_SCALE, CN_IMPULSE_SCALE,
  CN_A1X, CN_A1Y, CN_B1X, CN_B1Y, CN_BASE_SEP1, CN_NI1, CN_TNI1, CN_TI1, CN_NM1, CN_TM1,
  CN_A2X, CN_A2Y, CN_B2X, CN_B2Y, CN_BASE_SEP2, CN_NI2, CN_TNI2, CN_TI2, CN_NM2, CN_TM2,
  CN_RESTITUTION, CN_REL_VEL1, CN_REL_VEL2, CN_POINT_COUNT,
  CN_FIELD_COUNT
};

// Device pointers of a batch.
struct BatchView {
  WorldHeader* worlds;
  BodySim* sims;
  BodyState* states;
  Shape* shapes;
  ShapeGeom* geoms;
  Contact* contacts;
  int* freeIds;
  Joint* joints;
  uint64_t* pairs;      // new-pair scratch: sort key (hi) | local pair index (lo)
  int2* pairShapes;     // new-pair shapes (query order, unoriented)
  uint32_t* colorBits;
  int* overflow;        // overflow-colour contact ids in array order
  int* colorLists;      // per world at colorListOffset: colour arrays of colours 0..10
  float* cons;          // constraint SoA: field f of lane i at cons[f * consStride + i]
  long consStride;
  int* scratch;         // large-world sort scratch
  float4* sapBoxes;     // large-world sorted fat AABBs
  int8_t* ccolor;       // per contact slot: constraint-graph colour, -1 when none or free
  int* staticShapes;    // per world, the shape ids of its static proxies in the static tree's query order
  int* sweepOrder;      // per world, its shape ids in the last pair update's sweep order
  int* events;          // per world at 3 * contactOffset: begin contact ids [contactCap], end shape pairs [2 * contactCap]
  int* islandData;      // per world at islandOffset: the island region
  int4* contactLinks;   // per contact slot: island id, island prev, island next, solver set
  // broad-phase trees: per world three headers (static, kinematic, dynamic); each
  // tree's nodes and freed-id stack at its header's offset
  TreeHeader* treeHdrs;
  TreeNode* treeNodes;
  int* treeFreed;
  int* treeWork;        // per tree at treeWorkOffset: rebuild work arrays of trees beyond LDS
  int* treeRanks;       // per shape (shape offsets): its leaf's place in its tree's query order
  int worldCount;
};

}  // namespace f2d
