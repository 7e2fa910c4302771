// Host world store: the authoritative description of worlds while they are built,
// and the host mirror of device state after a step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "island.hpp"

namespace f2d {

struct HostWorld {
  WorldHeader hdr = {};
  std::vector<BodySim> sims;
  std::vector<BodyState> states;
  std::vector<Shape> shapes;
  std::vector<ShapeGeom> geoms;
  std::vector<Joint> joints;
  std::vector<uint16_t> jointGen;  // b2Joint::generation, raised at each creation in the slot
  std::vector<int> jointFree;      // the joint id pool's free list (b2IdPool, LIFO)
  std::vector<Contact> contacts;   // size = contact capacity after commit
  std::vector<int> freeIds;        // LIFO stack
  std::vector<int> overflow;
  std::vector<int> colorList[kOverflowIndex];  // colours 0..10: contact ids in array order
  std::vector<uint32_t> colorBits; // kOverflowIndex x colorWords
  std::vector<float> bodyMass, bodyInertia, shapeDensity;
  std::vector<uint16_t> bodyGen;
  std::vector<int> bodyFixedRotation;
  std::vector<std::string> bodyNames;  // b2Body name (31 characters), set through the flat ABI
  // broad-phase trees (static, kinematic, dynamic): headers, nodes, freed-id stacks
  TreeHeader tree[3] = {{-1, 0, 0, 0, 0, 0, 0, 0}, {-1, 0, 0, 0, 0, 0, 0, 0}, {-1, 0, 0, 0, 0, 0, 0, 0}};
  std::vector<TreeNode> tnodes[3];
  std::vector<int> tfreed[3];
  std::vector<int> treeRanks;      // per shape: its leaf's place in its tree's query order
  int createdSeq = 0;              // proxies created since the last step
  int contactCount = 0;
  int stepIndex = 0;
  float contactSpeed = 0.0f, inv_h = 0.0f;
  int colorWords = 0;
  bool hasCallbacks = false;
  std::vector<int> beginIds;       // contact begin-touch events of the last step (contact ids)
  std::vector<int> endPairs;       // contact end-touch events of the last step (shape pairs)
  bool eventsRetired = false;      // a paused step (dt = 0) since the last step: no events
  // island graph and solver sets (counters in hdr): records, id-pool free lists,
  // the awake set's island and body orders, the last finalize's body order
  std::vector<Island> islands;
  std::vector<int> islandFree, awakeIslands, setFree, awakeBodies, moveOrder;
  std::vector<int4> links;         // per contact slot: island id, prev, next, solver set
  int staticBodies = 0;            // static set size (static bodies' index in it)

  IslandCtx islandCtx();           // host view of the island data (valid until the vectors grow)
  TreeRef treeRef(int type, int extra = 0);  // host view of a tree with room for `extra` more nodes
  void applyPendingEnlarges();     // enlargements the device recorded but has not applied yet
  void computeTreeRanks();         // every proxy's rank in its tree's query order

  void ensureColorWords();
  int createBody(int type, Vec2 p, Rot q, Vec2 v, float w, float linDamp, float angDamp, float gravityScale,
                 float sleepThreshold, int flags, bool isAwake = true);
  int createShape(int body, const ShapeGeom& g, float density, float friction, float restitution, float rolling,
                  float tangentSpeed, uint64_t cat, uint64_t mask, int group, int flags, bool invokeContact,
                  bool updateMass);
  int createRevolute(int bodyA, int bodyB, Vec2 la, Vec2 lb, float refAngle, float targetAngle, bool enableSpring,
                     float hertz, float damping, bool enableLimit, float lower, float upper, bool enableMotor,
                     float maxMotorTorque, float motorSpeed, bool collideConnected);
  int createPrismatic(int bodyA, int bodyB, Vec2 la, Vec2 lb, Vec2 axis, float refAngle, float targetTranslation,
                      bool enableSpring, float hertz, float damping, bool enableLimit, float lower, float upper,
                      bool enableMotor, float maxMotorForce, float motorSpeed, bool collideConnected);
  int createWeld(int bodyA, int bodyB, Vec2 la, Vec2 lb, float refAngle, float linearHertz, float angularHertz,
                 float linearDamping, float angularDamping, bool collideConnected);
  int createWheel(int bodyA, int bodyB, Vec2 la, Vec2 lb, Vec2 axis, bool enableSpring, float hertz, float damping,
                  bool enableLimit, float lower, float upper, bool enableMotor, float maxMotorTorque,
                  float motorSpeed, bool collideConnected);
  int createDistance(int bodyA, int bodyB, Vec2 la, Vec2 lb, float length, bool enableSpring, float hertz,
                     float damping, bool enableLimit, float minLength, float maxLength, bool enableMotor,
                     float maxMotorForce, float motorSpeed, bool collideConnected);
  int createMotor(int bodyA, int bodyB, Vec2 linearOffset, float angularOffset, float maxForce, float maxTorque,
                  float correctionFactor, bool collideConnected);
  int createMouse(int bodyA, int bodyB, Vec2 target, float hertz, float damping, float maxForce,
                  bool collideConnected);
  int createFilter(int bodyA, int bodyB);
  int destroyJoint(int id);
  void updateBodyMass(int body);
};

struct DeviceBatch {
  BatchView view = {};
  std::vector<void*> allocations;
  size_t bytes = 0;
};

// Worlds of similar size launched together: each class gets dynamic LDS sized
// for its largest world and its own workgroup size, so small worlds are not
// limited to the occupancy the batch's largest world allows.
struct SizeClass {
  int limit = 0;              // max(bodies, shapes) of the class's worlds is at most this
  int count = 0;
  int maxBodies = 0, maxShapes = 0, maxJoints = 0;
  int* dList = nullptr;       // world ids, largest first
  LaunchConfig launch = {0, 0, 0, 0, 0};
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t tuneEv[6] = {};  // online lane tuning: start and end of each candidate's trial step
};

struct Batch {
  std::vector<HostWorld> worlds;
  DeviceBatch dev;
  bool committed = false;
  bool hostDirty = true;     // host holds changes the device lacks
  bool deviceNewer = false;  // device holds a step the host mirror lacks
  hipStream_t stream = nullptr;
  bool ownsStream = false;
  int* dWorldList = nullptr;
  int dWorldListCap = 0;
  int* dOrder = nullptr;     // every world, largest first: the launch order of full-batch steps
  std::vector<SizeClass> classes;
  std::vector<int> classOf;  // per world: its index in classes
  hipEvent_t forkEvent = nullptr;
  std::string lastError;
  float contactCapFactor = 6.0f;
  // flat array lengths of the committed batch (bulk host mirror copies)
  long nBodies = 0, nShapes = 0, nContacts = 0, nJoints = 0, nColorWords = 0, nIsland = 0, nTreeNodes = 1, nColorListInts = 1;
  LaunchConfig launch = {0, 0, 0, 0, 0};
  // query staging on the device (grown on demand, freed with the batch)
  Query* dQueries = nullptr;
  long dQueryCap = 0;
  QueryHit* dHits = nullptr;
  long dHitCap = 0;
  int* dCounts = nullptr;
  int* dGroup = nullptr;     // query grouping by world: the order, then per world a count and first slot
  long dGroupCap = 0;
  int forcedBlock = 0;       // f2dg_set_block before the batch is laid out: every class's size
  int tunePeriod = 0;        // online lane tuning: a trial round every tunePeriod steps (0: off)
  long tuneClock = 0;        // steps taken through f2dg_step since tuning was switched on
  long tuneNext = 0;         // the clock at which the next trial round is due
};

}  // namespace f2d
