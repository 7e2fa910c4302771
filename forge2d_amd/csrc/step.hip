// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Fused world step for gfx950: one workgroup advances one world through a whole
// b2World_Step (S/world.c:695-812) with workgroup barriers between the stages.
//
//   pairs     sort-and-sweep over fat AABBs (LDS bitonic sort of lower-x keys),
//             new pairs emitted with wavefront atomics, ordered like the
//             reference's move buffer, then contacts created serially
//             (S/broad_phase.c:160-474, S/contact.c:186-328)
//   collide   one contact per lane: overlap test, manifold, warm-start id
//             matching (S/world.c:357-445, S/contact.c:472-633)
//   state     contact begin/end/destroy in ascending contact id with persistent
//             greedy colouring (S/world.c:587-690, S/constraint_graph.c:68-212)
//   solve     prepare, then substeps x (integrate velocities, warm start, solve,
//             integrate positions, relax), restitution, store; overflow colour
//             first in each stage, then one colour per barrier, one constraint
//             per lane, body states resident in LDS (S/solver.c:929-1106)
//   finalize  one body per lane: transforms, AABBs, fat-AABB enlarge and the
//             next move buffer (S/solver.c:543-726, 1819-1911)
//
// Thousands of independent worlds are stepped by one launch (one workgroup
// each); a single world is the same launch with one workgroup.
#include <hip/hip_runtime.h>

#include "manifold.hpp"
#include "solver.hpp"
#include "joint.hpp"
#include "toi.hpp"
#include "island.hpp"
#include "query.hpp"
#include "tree_device.hpp"

namespace f2d {

constexpr int kMaxWaves = 16;  // largest workgroup: 1024 lanes
constexpr int kLongSweeps = 256;  // long sweeps a workgroup continues together (all of a 256-shape chunk)
constexpr int kLongRun = 32;      // candidates one lane tests before it hands a sweep over
// Colour table of a split solve at the front of the world's scratch: constraint
// counts and starts by colour, joint counts and starts by colour
constexpr int kWideColorCount = 0, kWideColorStart = 12, kWideJointCount = 25, kWideJointStart = 37;
constexpr int kWideTableInts = 64;

struct SharedState {
  int colorCount[kGraphColorCount];
  int colorStart[kGraphColorCount + 1];
  int colorFill[kGraphColorCount];
  int jointColorCount[kGraphColorCount];
  int jointColorStart[kGraphColorCount + 1];
  int jointColorFill[kGraphColorCount];
  int waveSums[kMaxWaves];
  int counter;
  int fastCount;
  int changedCount;
  int error;
  int pairCount;
  int noAwake;  // no awake body: b2Solve returns after merging islands
};

// Sweeps of shapes whose x-range covers many others (ground and ceiling segments),
// continued by the whole workgroup: shape, next candidate (collide and sweep
// kernels only, so the solve kernel's LDS stays small)
struct LongSweeps {
  int count;
  int shape[kLongSweeps];
  int next[kLongSweeps];
};

// ---------------------------------------------------------------------------- block helpers

// Ordered block-wide compaction step: returns this thread's slot among flagged
// threads of the block and writes the block total.
template <int BLOCK>
__device__ __forceinline__ int blockCompact(bool flag, int* waveSums, int* total) {
  constexpr int kWaves = BLOCK / 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  unsigned long long mask = __ballot(flag);
  unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int prefix = __popcll(mask & lt);
  if (lane == 0) waveSums[wave] = __popcll(mask);
  __syncthreads();
  int base = 0, sum = 0;
  for (int w = 0; w < kWaves; ++w) {
    int s = waveSums[w];
    if (w < wave) base += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return base + prefix;
}

template <int BLOCK, typename K, typename V>
__device__ void bitonicSort(K* key, V* val, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += BLOCK) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          K a = key[i], b = key[ixj];
          if ((a > b) == up) {
            key[i] = b;
            key[ixj] = a;
            V t = val[i];
            val[i] = val[ixj];
            val[ixj] = t;
          }
        }
      }
      __syncthreads();
    }
  }
}

__device__ __forceinline__ uint32_t orderedKey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ int nextPow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// b2ShouldShapesCollide, S/shape.h:122-130
__device__ __forceinline__ bool shouldShapesCollide(const Shape& a, const Shape& b) {
  if (a.group == b.group && a.group != 0) return a.group > 0;
  uint64_t maskA = (uint64_t)a.maskLo | ((uint64_t)a.maskHi << 32);
  uint64_t catA = (uint64_t)a.catLo | ((uint64_t)a.catHi << 32);
  uint64_t maskB = (uint64_t)b.maskLo | ((uint64_t)b.maskHi << 32);
  uint64_t catB = (uint64_t)b.catLo | ((uint64_t)b.catHi << 32);
  return (maskA & catB) != 0 && (catA & maskB) != 0;
}

// b2ShouldBodiesCollide, S/body.c:1861-1897 (joint part; the type test is done by the caller)
__device__ __forceinline__ bool shouldBodiesCollide(const BodySim* sims, const Joint* joints, int bodyA, int bodyB) {
  int jointKey, other;
  if (sims[bodyA].jointCount < sims[bodyB].jointCount) {
    jointKey = sims[bodyA].headJointKey;
    other = bodyB;
  } else {
    jointKey = sims[bodyB].headJointKey;
    other = bodyA;
  }
  for (int guard = sims[bodyA].jointCount + sims[bodyB].jointCount; jointKey >= 0 && guard >= 0; --guard) {
    int jointId = jointKey >> 1;
    int edge = jointKey & 1;
    const Joint& j = joints[jointId];
    int otherBody = edge == 0 ? j.bodyB : j.bodyA;
    if (j.collideConnected == 0 && otherBody == other) return false;
    jointKey = edge == 0 ? j.nextA : j.nextB;
  }
  return true;
}

// Does a contact between shapes s and t exist? Walks the shorter body contact list
// (the reference keeps the same information in its pair set, S/broad_phase.c:221-226).
__device__ __forceinline__ bool contactExists(const BodySim* sims, const Contact* contacts, int bodyS, int bodyT,
                                              int s, int t) {
  int body = sims[bodyS].contactCount <= sims[bodyT].contactCount ? bodyS : bodyT;
  int key = sims[body].headContactKey;
  // the walk is bounded by the list length; a longer walk means a corrupted list
  for (int guard = sims[body].contactCount; key >= 0 && guard >= 0; --guard) {
    const Contact& c = contacts[key >> 1];
    if ((c.shapeA == s && c.shapeB == t) || (c.shapeA == t && c.shapeB == s)) return true;
    key = (key & 1) == 0 ? c.nextA : c.nextB;
  }
  return false;
}

// ---------------------------------------------------------------------------- phase: pairs

struct WorldPtrs {
  WorldHeader* hdr;
  BodySim* sims;
  BodyState* states;
  Shape* shapes;
  ShapeGeom* geoms;
  Contact* contacts;
  int* freeIds;
  Joint* joints;
  uint64_t* pairKeys;
  int2* pairShapes;
  uint32_t* colorBits;
  int* overflow;
  int* colorLists;  // counts, then colour arrays of colours 0..10 (kColorListHeader layout)
  Cons cons;
  int* scratch;
  float4* sapBoxes;
  int8_t* ccolor;
  int* events;
  int4* links;      // contact island links and solver set
  int* sweepOrder;
  const int* treeRanks;  // per shape: its leaf's place in its tree's query order
  IslandCtx isl;
};

__device__ WorldPtrs worldPtrs(const BatchView& b, WorldHeader* h) {
  WorldPtrs w;
  w.hdr = h;
  w.sims = b.sims + h->bodyOffset;
  w.states = b.states + h->bodyOffset;
  w.shapes = b.shapes + h->shapeOffset;
  w.geoms = b.geoms + h->shapeOffset;
  w.contacts = b.contacts + h->contactOffset;
  w.freeIds = b.freeIds + h->freeOffset;
  w.joints = b.joints + h->jointOffset;
  w.pairKeys = b.pairs + h->pairOffset;
  w.pairShapes = b.pairShapes + h->pairOffset;
  w.colorBits = b.colorBits + h->colorOffset;
  w.overflow = b.overflow + h->overflowOffset;
  w.colorLists = b.colorLists + h->colorListOffset;
  w.cons.f = b.cons + h->constraintOffset;
  w.cons.stride = b.consStride;
  w.scratch = b.scratch + h->scratchOffset;
  w.ccolor = b.ccolor + h->contactOffset;
  w.events = b.events + 3L * h->contactOffset;
  w.sapBoxes = b.sapBoxes + h->shapeOffset;
  w.sweepOrder = b.sweepOrder + h->shapeOffset;
  w.treeRanks = b.treeRanks + h->shapeOffset;
  w.links = b.contactLinks + h->contactOffset;
  w.isl = islandCtx(h, b.islandData + h->islandOffset, w.sims, w.contacts, w.links, w.joints);
  return w;
}

// Existing contacts of a world as an LDS hash set of shape-pair keys (open
// addressing, linear probing): the pair test asks it instead of walking the two
// bodies' contact lists in HBM. Keys are (min << 16 | max) of the shape ids.
__device__ __forceinline__ uint32_t pairSetKey(int s, int t) {
  return s < t ? ((uint32_t)s << 16) | (uint32_t)t : ((uint32_t)t << 16) | (uint32_t)s;
}
__device__ __forceinline__ uint32_t pairSetSlot(uint32_t key, uint32_t mask) {
  return (key * 2654435761u) & mask;
}
__device__ __forceinline__ bool pairSetHas(const uint32_t* table, uint32_t mask, uint32_t key) {
  for (uint32_t i = pairSetSlot(key, mask), n = 0; n <= mask; i = (i + 1) & mask, ++n) {
    uint32_t v = table[i];
    if (v == key) return true;
    if (v == 0xFFFFFFFFu) return false;
  }
  return false;
}

// What the pair test reads of a shape: body, proxy key, flags (plus whether the
// body has joints), filter. Cached in LDS, sorted like the sweep, for worlds that fit.
constexpr int kPairBodyHasJoints = 1 << 29;
struct alignas(16) PairInfo {
  int body, proxyKey, flags, group;
  uint32_t catLo, catHi, maskLo, maskHi;
};

__device__ __forceinline__ PairInfo pairInfo(const WorldPtrs& w, int s) {
  const Shape& S = w.shapes[s];
  PairInfo p;
  p.body = S.body;
  p.proxyKey = S.proxyKey;
  p.flags = S.flags | (w.sims[S.body].jointCount > 0 ? kPairBodyHasJoints : 0);
  p.group = S.group;
  p.catLo = S.catLo;
  p.catHi = S.catHi;
  p.maskLo = S.maskLo;
  p.maskHi = S.maskHi;
  return p;
}

// b2ShouldShapesCollide, S/shape.h:122-130
__device__ __forceinline__ bool shouldPairCollide(const PairInfo& a, const PairInfo& b) {
  if (a.group == b.group && a.group != 0) return a.group > 0;
  uint64_t maskA = (uint64_t)a.maskLo | ((uint64_t)a.maskHi << 32);
  uint64_t catA = (uint64_t)a.catLo | ((uint64_t)a.catHi << 32);
  uint64_t maskB = (uint64_t)b.maskLo | ((uint64_t)b.maskHi << 32);
  uint64_t catB = (uint64_t)b.catLo | ((uint64_t)b.catHi << 32);
  return (maskA & catB) != 0 && (catA & maskB) != 0;
}

// Evaluate one overlapping candidate (s, t) of the sweep. Emits a new pair when the
// reference's pair query would: at least one proxy moved, tree rules, filters, no
// existing contact. The emitter is the proxy whose query reports the pair.
__device__ void considerPair(const WorldPtrs& w, int* pairCount, int s, int t, const PairInfo& S,
                             const PairInfo& T, const uint32_t* pairSet, uint32_t pairMask) {
  int bs = S.body, bt = T.body;
  if (bs == bt) return;
  // the proxy type is the body type (S/broad_phase.h:18-21)
  int typeS = S.proxyKey & 3, typeT = T.proxyKey & 3;
  if (typeS != kDynamic && typeT != kDynamic) return;
  bool movedS = (S.flags & SF_MOVED) != 0, movedT = (T.flags & SF_MOVED) != 0;
  int emitter;
  if (typeS == kDynamic && typeT == kDynamic) {
    if (movedS && movedT) {
      emitter = S.proxyKey < T.proxyKey ? s : t;
    } else if (movedS) {
      emitter = s;
    } else if (movedT) {
      emitter = t;
    } else {
      return;
    }
  } else {
    int d = typeS == kDynamic ? s : t;
    int x = typeS == kDynamic ? t : s;
    bool movedD = typeS == kDynamic ? movedS : movedT;
    bool movedX = typeS == kDynamic ? movedT : movedS;
    if (movedD) {
      emitter = d;
    } else if (movedX) {
      emitter = x;
    } else {
      return;
    }
  }
  const int partner = emitter == s ? t : s;
  if ((S.flags | T.flags) & SF_SENSOR) return;
  if (!shouldPairCollide(S, T)) return;
  // b2ShouldBodiesCollide only vetoes pairs whose bodies share a joint
  if ((S.flags & kPairBodyHasJoints) && (T.flags & kPairBodyHasJoints) &&
      !shouldBodiesCollide(w.sims, w.joints, bs, bt))
    return;
  if (pairSet ? pairSetHas(pairSet, pairMask, pairSetKey(s, t)) : contactExists(w.sims, w.contacts, bs, bt, s, t))
    return;
  const PairInfo& P = partner == s ? S : T;
  // Order within one move result: dynamic-tree partners, then static, then kinematic
  // (the query order reversed by the list prepend, S/broad_phase.c:345-371, 285-287).
  // Within one tree the query visits partners in its leaf order (child2 first) and the
  // prepend reverses it (S/dynamic_tree.c:1114-1170, S/broad_phase.c:285-287).
  int ptype = P.proxyKey & 3;
  int treeOrder = ptype == kDynamic ? 0 : (ptype == kStatic ? 1 : 2);
  uint32_t partnerOrder = 0x3FFFFFFFu - (uint32_t)w.treeRanks[partner];
  uint64_t key = ((uint64_t)(uint32_t)w.shapes[emitter].moveRank << 32) | ((uint64_t)treeOrder << 30) | partnerOrder;
  int idx = atomicAdd(pairCount, 1);
  if (idx < w.hdr->pairCap) {
    w.pairKeys[idx] = key;
    w.pairShapes[idx] = make_int2(emitter, partner);
  }
}

// b2CreateContact (S/contact.c:186-328) for the sorted new pairs (pv: pair indices in
// creation order), with the result of a serial loop over them: ids are the LIFO
// pool's pops in creation order (free list top first, then fresh ids), so each pair's
// id follows from its rank among the pairs that get a contact; the records are filled
// in parallel and only the body contact-list prepends run serially, in pair order.
// pk is reused for the oriented shapes and pv for the ids.
template <int BLOCK>
__device__ void createContacts(const WorldPtrs& w, SharedState& sh, uint64_t* pk, int* pv, int count) {
  WorldHeader* h = w.hdr;
  const int freeCount = h->freeCount, fresh = h->contactCount, cap = h->contactCap;
  int created = 0;
  for (int base = 0; base < count; base += BLOCK) {
    const int i = base + threadIdx.x;
    int a = -1, bsh = -1;
    bool valid = false;
    if (i < count) {
      int2 pr = w.pairShapes[pv[i]];
      int e = pr.x, p = pr.y;
      // orientation by proxy key, S/broad_phase.c:228-237, then the primary type order
      if (w.shapes[p].proxyKey < w.shapes[e].proxyKey) {
        a = p;
        bsh = e;
      } else {
        a = e;
        bsh = p;
      }
      int ta = w.geoms[a].type, tb = w.geoms[bsh].type;
      valid = hasManifoldFcn(ta, tb);
      if (valid && !isPrimary(ta, tb)) {
        int t = a;
        a = bsh;
        bsh = t;
      }
    }
    int total;
    const int rank = created + blockCompact<BLOCK>(valid, sh.waveSums, &total);
    if (i < count) {
      int id = -1;
      if (valid) {
        if (rank < freeCount) {
          id = w.freeIds[freeCount - 1 - rank];
        } else if (fresh + (rank - freeCount) < cap) {
          id = fresh + (rank - freeCount);
        } else {
          atomicOr(&h->error, ERR_CONTACT_CAP);
        }
      }
      pk[i] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)bsh;
      pv[i] = id;
      if (id >= 0) {
        const Shape& A = w.shapes[a];
        const Shape& B = w.shapes[bsh];
        Contact& c = w.contacts[id];
        c.shapeA = a;
        c.shapeB = bsh;
        c.flags = ((A.flags | B.flags) & SF_CONTACT_EVENTS) ? CF_CONTACT_EVENTS : 0;
        c.simFlags = ((A.flags | B.flags) & SF_PRESOLVE) ? SIM_PRESOLVE : 0;
        c.color = -1;
        c.localIndex = -1;
        w.ccolor[id] = -1;
        c.bodyA = A.body;
        c.bodyB = B.body;
        zeroManifold(c.m);
        c.friction = sqrtf(A.friction * B.friction);
        c.restitution = maxf(A.restitution, B.restitution);
        c.rollingResistance = 0.0f;
        c.tangentSpeed = 0.0f;
        // the awake set when either body is awake, else the disabled set (S/contact.c:213-224)
        int set = (w.sims[A.body].setIndex == kAwakeSet || w.sims[B.body].setIndex == kAwakeSet) ? kAwakeSet
                                                                                                   : kDisabledSet;
        if (set != kAwakeSet) c.flags |= CF_OFF_AWAKE;
        w.links[id] = make_int4(-1, -1, -1, set);
      }
    }
    created += total;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int used = min(created, freeCount + (cap - fresh));
    h->freeCount = freeCount - min(used, freeCount);
    h->contactCount = fresh + max(used - freeCount, 0);
    // body contact lists: prepend each new contact to both bodies' lists in creation order
    for (int i = 0; i < count; ++i) {
      const int id = pv[i];
      if (id < 0) continue;
      Contact& c = w.contacts[id];
      for (int edge = 0; edge < 2; ++edge) {
        BodySim& body = w.sims[edge == 0 ? c.bodyA : c.bodyB];
        const int key = (id << 1) | edge;
        const int head = body.headContactKey;
        if (edge == 0) {
          c.prevA = -1;
          c.nextA = head;
        } else {
          c.prevB = -1;
          c.nextB = head;
        }
        if (head >= 0) {
          Contact& hc = w.contacts[head >> 1];
          if ((head & 1) == 0) hc.prevA = key; else hc.prevB = key;
        }
        body.headContactKey = key;
        body.contactCount += 1;
      }
    }
  }
  __syncthreads();
}

// The sweep of sorted shape i: every later shape j (from jBegin, step jStep) whose
// lower x lies within its box, tested on the fat boxes, then the reference's pair
// rules (considerPair). With jStep 1 it stops after maxRun candidates and returns
// the next j (the rest of a long sweep, continued by the whole workgroup), else -1.
__device__ __forceinline__ int sweepShape(const WorldPtrs& w, const float4* boxes, const int* order,
                                          const PairInfo* info, int i, int jBegin, int jStep, int n, int maxRun,
                                          int* pairCount, const uint32_t* pairSet, uint32_t pairMask) {
  float4 a = boxes[i];
  int s = order[i];
  const PairInfo S = info ? info[i] : pairInfo(w, s);
  int run = 0;
  for (int j = jBegin; j < n; j += jStep) {
    if (run == maxRun) return j;
    ++run;
    float4 bx = boxes[j];
    if (bx.x > a.z) break;
    // b2AABB_Overlaps on the two fat boxes (x already known to overlap on one side)
    bool overlap = !(bx.x > a.z || bx.y > a.w || a.x > bx.z || a.y > bx.w);
    if (overlap) {
      const int t = order[j];
      const PairInfo T = info ? info[j] : pairInfo(w, t);
      considerPair(w, pairCount, s, t, S, T, pairSet, pairMask);
    }
  }
  return -1;
}

// Every sweep of a workgroup's shapes [first, first + count): each lane sweeps its
// shapes up to kLongRun candidates; longer sweeps are continued by all lanes, one
// after another, each lane testing every BLOCK-th candidate (the sorted boxes end
// a lane's part at the first box past the shape's upper x). Every (i, j) candidate
// is tested exactly once; the pairs' order is fixed later by their sort.
template <int BLOCK>
__device__ void sweepShapes(const WorldPtrs& w, LongSweeps& sh, const float4* boxes, const int* order,
                            const PairInfo* info, int first, int count, int n, int* pairCount,
                            const uint32_t* pairSet, uint32_t pairMask) {
  if (threadIdx.x == 0) sh.count = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < count; k += BLOCK) {
    const int i = first + k;
    const int next = sweepShape(w, boxes, order, info, i, i + 1, 1, n, kLongRun, pairCount, pairSet, pairMask);
    if (next >= 0) {
      const int slot = atomicAdd(&sh.count, 1);
      if (slot < kLongSweeps) {
        sh.shape[slot] = i;
        sh.next[slot] = next;
      } else {
        sweepShape(w, boxes, order, info, i, next, 1, n, -1, pairCount, pairSet, pairMask);
      }
    }
  }
  __syncthreads();
  const int nLong = min(sh.count, kLongSweeps);
  for (int k = 0; k < nLong; ++k) {
    const int i = sh.shape[k];
    sweepShape(w, boxes, order, info, i, sh.next[k] + threadIdx.x, BLOCK, n, -1, pairCount, pairSet, pairMask);
  }
  __syncthreads();
}

// Pair update of one world (b2UpdateBroadPhasePairs). mode 0: the whole update in
// this workgroup. A world whose sort arrays do not fit LDS can be split over three
// launches: mode 1 sorts (worlds that fit LDS run the whole update here), the sweep
// launch (sweepKernel) spreads the sweep over many workgroups, and mode 2 creates the
// pairs it found, in the same order.
template <int BLOCK>
__device__ void pairsPhase(const WorldPtrs& w, SharedState& sh, char* lds, int ldsShapes, int ldsPairs, int ldsBytes,
                           int mode) {
  WorldHeader* h = w.hdr;
  const int n = h->shapeCount;
  if (threadIdx.x == 0) sh.pairCount = 0;
  int count = 0;
  if (mode == 2) {
    if (h->sweepCount == 0) return;
    count = h->sweepPairs;
    __syncthreads();
    if (threadIdx.x == 0) h->sweepCount = 0;
  } else {
  if (mode == 1 && threadIdx.x == 0) h->sweepCount = 0;
  // any moved proxy at all?
  int anyMoved = 0;
  for (int i = threadIdx.x; i < n; i += BLOCK) anyMoved |= (w.shapes[i].flags & SF_MOVED) ? 1 : 0;
  anyMoved = __syncthreads_or(anyMoved);
  if (!anyMoved) return;

  // LDS layout by the world's padded shape count: sort keys (8 B), sorted fat boxes
  // (16 B) and shape ids (4 B) per shape, then the pair-test cache (32 B per shape)
  // and the existing-pair hash set when they fit
  const int np = nextPow2(max(n, 4));
  const bool inLds = 28L * np <= (long)ldsBytes;
  (void)ldsShapes;
  uint64_t* keys = inLds ? reinterpret_cast<uint64_t*>(lds) : reinterpret_cast<uint64_t*>(w.scratch);
  // keys (lower fat x, shape id) laid out in the last step's sweep order, which is
  // nearly sorted when bodies move little per step
  int* sweepOrder = w.sweepOrder;
  for (int i = threadIdx.x; i < np; i += BLOCK) {
    uint64_t k = ~0ull;
    if (i < n) {
      const int s = sweepOrder[i];
      k = ((uint64_t)orderedKey(w.shapes[s].fat.lo.x) << 32) | (uint32_t)s;
    }
    keys[i] = k;
  }
  __syncthreads();
  // odd-even transposition passes while they still swap (a few for coherent
  // motion); a bounded number of them, then a full bitonic sort. The keys are
  // distinct, so the sorted order is the same either way.
  bool sorted = false;
  const int maxPasses = inLds ? 16 : 64;  // a bitonic sort in HBM costs about 200 passes
  for (int pass = 0; pass < maxPasses && !sorted; ++pass) {
    int swapped = 0;
    for (int parity = 0; parity < 2; ++parity) {
      for (int i = 2 * threadIdx.x + parity; i + 1 < n; i += 2 * BLOCK) {
        uint64_t a = keys[i], b = keys[i + 1];
        if (a > b) {
          keys[i] = b;
          keys[i + 1] = a;
          swapped = 1;
        }
      }
      __syncthreads();
    }
    sorted = __syncthreads_or(swapped) == 0;
  }
  if (!sorted) {
    for (int k = 2; k <= np; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < np; i += BLOCK) {
          int ixj = i ^ j;
          if (ixj > i) {
            bool up = (i & k) == 0;
            uint64_t a = keys[i], b = keys[ixj];
            if ((a > b) == up) {
              keys[i] = b;
              keys[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  // sorted boxes and indices
  float4* boxes = inLds ? reinterpret_cast<float4*>(lds + np * 8) : w.sapBoxes;
  int* order = inLds ? reinterpret_cast<int*>(lds + np * 24) : (w.scratch + 2 * np);
  PairInfo* info = (inLds && 60L * np <= (long)ldsBytes) ? reinterpret_cast<PairInfo*>(lds + np * 28) : nullptr;
  for (int i = threadIdx.x; i < n; i += BLOCK) {
    int idx = (int)(uint32_t)keys[i];
    AABB f = w.shapes[idx].fat;
    boxes[i] = make_float4(f.lo.x, f.lo.y, f.hi.x, f.hi.y);
    order[i] = idx;
    sweepOrder[i] = idx;
    if (info) info[i] = pairInfo(w, idx);
  }
  __syncthreads();
  if (mode == 1 && !inLds) {
    // the sweep runs in its own launch over many workgroups
    if (threadIdx.x == 0) {
      h->sweepPairs = 0;
      h->sweepCount = n;
    }
    return;
  }
  // the existing pairs as an LDS hash set when it fits behind the sort arrays
  uint32_t* pairSet = nullptr;
  uint32_t pairMask = 0;
  if (inLds && n < 65536) {
    const int live = h->contactCount - h->freeCount;
    const int tsize = nextPow2(max(2 * live, 64));
    const long base = info ? 60L * np : 28L * np;
    if (base + 4L * tsize <= (long)ldsBytes) {
      pairSet = reinterpret_cast<uint32_t*>(lds + base);
      pairMask = (uint32_t)tsize - 1u;
      for (int i = threadIdx.x; i < tsize; i += BLOCK) pairSet[i] = 0xFFFFFFFFu;
      __syncthreads();
      for (int i = threadIdx.x; i < h->contactCount; i += BLOCK) {
        const Contact& c = w.contacts[i];
        if (c.shapeA < 0) continue;
        const uint32_t key = pairSetKey(c.shapeA, c.shapeB);
        uint32_t k = pairSetSlot(key, pairMask);
        for (uint32_t probe = 0; probe <= pairMask; ++probe, k = (k + 1) & pairMask) {
          if (atomicCAS(&pairSet[k], 0xFFFFFFFFu, key) == 0xFFFFFFFFu) break;
        }
      }
      __syncthreads();
    }
  }
  __shared__ LongSweeps longSweeps;
  sweepShapes<BLOCK>(w, longSweeps, boxes, order, info, 0, n, n, &sh.pairCount, pairSet, pairMask);
  count = sh.pairCount;
  }
  if (count > h->pairCap) {
    if (threadIdx.x == 0) h->error |= ERR_PAIR_CAP;
    count = h->pairCap;
  }
  if (count > 0) {
    // sort by move order key; values are pair indices
    const int pp = nextPow2(count);
    const bool pairsInLds = pp <= ldsPairs;
    uint64_t* pk = pairsInLds ? reinterpret_cast<uint64_t*>(lds) : reinterpret_cast<uint64_t*>(w.scratch);
    int* pv = pairsInLds ? reinterpret_cast<int*>(lds + ldsPairs * 8) : (w.scratch + 2 * pp);
    __syncthreads();
    for (int i = threadIdx.x; i < pp; i += BLOCK) {
      pk[i] = i < count ? w.pairKeys[i] : ~0ull;
      pv[i] = i;
    }
    __syncthreads();
    bitonicSort<BLOCK>(pk, pv, pp);
    createContacts<BLOCK>(w, sh, pk, pv, count);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += BLOCK) w.shapes[i].flags &= ~SF_MOVED;
  __syncthreads();
}

// ---------------------------------------------------------------------------- phase: collide

// b2UpdateContact, S/contact.c:472-633 (no pre-solve callback on the device)
__device__ void updateContact(const WorldPtrs& w, Contact& c) {
  const Shape& A = w.shapes[c.shapeA];
  const Shape& B = w.shapes[c.shapeB];
  const BodySim& simA = w.sims[A.body];
  const BodySim& simB = w.sims[B.body];
  Xf xfA = simA.xf, xfB = simB.xf;
  Vec2 centerOffsetA = rotate(xfA.q, simA.localCenter);
  Vec2 centerOffsetB = rotate(xfB.q, simB.localCenter);
  bool wasTouching = (c.simFlags & SIM_TOUCHING) != 0;
  Manifold old = c.m;
  Manifold m;
  computeManifold(m, w.geoms[c.shapeA], xfA, w.geoms[c.shapeB], xfB);
  c.friction = sqrtf(A.friction * B.friction);
  c.restitution = maxf(A.restitution, B.restitution);
  if (A.rollingResistance > 0.0f || B.rollingResistance > 0.0f) {
    float radiusA = shapeRadius(w.geoms[c.shapeA]);
    float radiusB = shapeRadius(w.geoms[c.shapeB]);
    float maxRadius = maxf(radiusA, radiusB);
    c.rollingResistance = maxf(A.rollingResistance, B.rollingResistance) * maxRadius;
  } else {
    c.rollingResistance = 0.0f;
  }
  c.tangentSpeed = A.tangentSpeed + B.tangentSpeed;
  int pointCount = m.pointCount;
  bool touching = pointCount > 0;
  if (touching && ((A.flags | B.flags) & SF_HIT_EVENTS)) {
    c.simFlags |= SIM_HIT_EVENT;
  } else {
    c.simFlags &= ~SIM_HIT_EVENT;
  }
  if (pointCount > 0) m.rollingImpulse = old.rollingImpulse;
  // The old points' impulses as scalars: a matched old point hands its impulses
  // over once and keeps its id, so a second new point with the same id matches it
  // again and takes zeros (S/contact.c:568-583).
  const int oldCount = old.pointCount;
  const int oldId0 = old.points[0].id & 0xFFFF, oldId1 = old.points[1].id & 0xFFFF;
  float oldN0 = old.points[0].normalImpulse, oldT0 = old.points[0].tangentImpulse;
  float oldN1 = old.points[1].normalImpulse, oldT1 = old.points[1].tangentImpulse;
  F2D_UNROLL
  for (int i = 0; i < 2; ++i) {
    if (i >= pointCount) break;
    ManifoldPoint& mp2 = m.points[i];
    mp2.anchorA = sub(mp2.anchorA, centerOffsetA);
    mp2.anchorB = sub(mp2.anchorB, centerOffsetB);
    mp2.normalImpulse = 0.0f;
    mp2.tangentImpulse = 0.0f;
    mp2.totalNormalImpulse = 0.0f;
    mp2.normalVelocity = 0.0f;
    int id2 = mp2.id & 0xFFFF;
    mp2.id = id2;  // persisted = false
    if (oldCount > 0 && oldId0 == id2) {
      mp2.normalImpulse = oldN0;
      mp2.tangentImpulse = oldT0;
      mp2.id = id2 | 0x10000;
      oldN0 = 0.0f;
      oldT0 = 0.0f;
    } else if (oldCount > 1 && oldId1 == id2) {
      mp2.normalImpulse = oldN1;
      mp2.tangentImpulse = oldT1;
      mp2.id = id2 | 0x10000;
      oldN1 = 0.0f;
      oldT1 = 0.0f;
    }
  }
  if (touching) {
    c.simFlags |= SIM_TOUCHING;
  } else {
    c.simFlags &= ~SIM_TOUCHING;
  }
  if (touching && !wasTouching) {
    c.simFlags |= SIM_STARTED;
  } else if (!touching && wasTouching) {
    c.simFlags |= SIM_STOPPED;
  }
  c.m = m;
}

// The narrow phase of contact slot i: awake contacts whose fat boxes stopped
// overlapping are flagged disjoint, the others get a new manifold.
__device__ __forceinline__ void collideContact(const WorldPtrs& w, int i) {
  Contact& c = w.contacts[i];
  if (c.shapeA < 0 || (c.flags & CF_OFF_AWAKE)) return;  // only the awake set collides (S/world.c:496-521)
  AABB fa = w.shapes[c.shapeA].fat, fb = w.shapes[c.shapeB].fat;
  if (!aabbOverlaps(fa, fb)) {
    c.simFlags |= SIM_DISJOINT;
    c.simFlags &= ~SIM_TOUCHING;
  } else {
    updateContact(w, c);
  }
}

// contactsDone: a narrow-phase launch already updated the contacts
template <int BLOCK>
__device__ void collidePhase(const WorldPtrs& w, SharedState& sh, bool contactsDone) {
  WorldHeader* h = w.hdr;
  const int n = h->contactCount;
  if (!contactsDone) {
    for (int i = threadIdx.x; i < n; i += BLOCK) collideContact(w, i);
  }
  __syncthreads();
  // ordered list of contacts whose state changed (ascending contact id)
  if (threadIdx.x == 0) sh.changedCount = 0;
  __syncthreads();
  int* list = w.scratch;
  for (int base = 0; base < n; base += BLOCK) {
    int i = base + threadIdx.x;
    bool f = false;
    if (i < n) {
      const Contact& c = w.contacts[i];
      f = c.shapeA >= 0 && (c.flags & CF_OFF_AWAKE) == 0 &&
          (c.simFlags & (SIM_DISJOINT | SIM_STARTED | SIM_STOPPED)) != 0;
    }
    int total;
    int slot = blockCompact<BLOCK>(f, sh.waveSums, &total);
    if (f) list[sh.changedCount + slot] = i;
    __syncthreads();
    if (threadIdx.x == 0) sh.changedCount += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------- phase: contact state + colouring

__device__ __forceinline__ uint32_t* colorSet(const WorldPtrs& w, int color) {
  return w.colorBits + (long)color * w.hdr->colorWords;
}
__device__ __forceinline__ bool getBit(const uint32_t* set, int i) { return (set[i >> 5] >> (i & 31)) & 1u; }
__device__ __forceinline__ void setBit(uint32_t* set, int i) { set[i >> 5] |= 1u << (i & 31); }
__device__ __forceinline__ void clearBit(uint32_t* set, int i) { set[i >> 5] &= ~(1u << (i & 31)); }

// b2AddContactToGraph colour choice, S/constraint_graph.c:68-135: the first colour
// (from colour 1 when a body is static) that holds neither non-static body, or the
// overflow. The bodies' bits of every colour are read at once (one memory round
// trip) into a mask of the colours that hold either.
__device__ void addContactToGraphBodies(const WorldPtrs& w, int id, int bodyA, int bodyB, bool staticA,
                                        bool staticB) {
  Contact& c = w.contacts[id];
  int color = kOverflowIndex;
  if (!staticA || !staticB) {
    const long stride = w.hdr->colorWords;
    const uint32_t* wordA = w.colorBits + (bodyA >> 5);
    const uint32_t* wordB = w.colorBits + (bodyB >> 5);
    uint32_t busy = 0u;
    F2D_UNROLL
    for (int i = 0; i < kOverflowIndex; ++i) {
      const uint32_t a = staticA ? 0u : (wordA[i * stride] >> (bodyA & 31));
      const uint32_t b = staticB ? 0u : (wordB[i * stride] >> (bodyB & 31));
      busy |= ((a | b) & 1u) << i;
    }
    if (staticA || staticB) busy |= 1u;  // colour 0 is for dynamic pairs only
    const uint32_t freeColors = ~busy & ((1u << kOverflowIndex) - 1u);
    if (freeColors != 0u) {
      color = __builtin_ctz(freeColors);
      uint32_t* set = colorSet(w, color);
      if (!staticA) setBit(set, bodyA);
      if (!staticB) setBit(set, bodyB);
    }
  }
  c.color = color;
  w.ccolor[id] = (int8_t)color;
  if (color == kOverflowIndex) {
    WorldHeader* h = w.hdr;
    if (h->overflowCount >= h->contactCap) {
      h->error |= ERR_OVERFLOW_CAP;
      c.localIndex = -1;
      return;
    }
    c.localIndex = h->overflowCount;
    w.overflow[h->overflowCount++] = id;
  } else {
    // appended to the colour's array (b2ContactSimArray_Add, S/constraint_graph.c:160-175)
    const int n = w.colorLists[color]++;
    w.colorLists[kColorListHeader + (long)color * w.hdr->bodyCount + n] = id;
    c.localIndex = n;
  }
}

__device__ void addContactToGraph(const WorldPtrs& w, int id) {
  const Contact& c = w.contacts[id];
  const int bodyA = c.bodyA, bodyB = c.bodyB;
  addContactToGraphBodies(w, id, bodyA, bodyB, w.sims[bodyA].type == kStatic, w.sims[bodyB].type == kStatic);
}

// b2RemoveContactFromGraph, S/constraint_graph.c:184-212
__device__ void removeContactFromGraph(const WorldPtrs& w, int id) {
  Contact& c = w.contacts[id];
  int color = c.color;
  if (color < 0) return;
  if (color != kOverflowIndex) {
    uint32_t* set = colorSet(w, color);
    clearBit(set, c.bodyA);
    clearBit(set, c.bodyB);
    // remove-swap from the colour's array (S/constraint_graph.c:198-209)
    int* list = w.colorLists + kColorListHeader + (long)color * w.hdr->bodyCount;
    const int idx = c.localIndex, last = --w.colorLists[color];
    if (idx != last && idx >= 0) {
      const int moved = list[last];
      list[idx] = moved;
      w.contacts[moved].localIndex = idx;
    }
  } else {
    WorldHeader* h = w.hdr;
    int idx = c.localIndex;
    int last = --h->overflowCount;
    if (idx != last && idx >= 0) {
      int moved = w.overflow[last];
      w.overflow[idx] = moved;
      w.contacts[moved].localIndex = idx;
    }
  }
  c.color = -1;
  c.localIndex = -1;
  w.ccolor[id] = -1;
}

// b2DestroyContact, S/contact.c:337-454 (no island or event bookkeeping on the device)
__device__ void destroyContact(const WorldPtrs& w, int id) {
  Contact& c = w.contacts[id];
  // remove from body A list
  {
    if (c.prevA >= 0) {
      Contact& pc = w.contacts[c.prevA >> 1];
      if ((c.prevA & 1) == 0) pc.nextA = c.nextA; else pc.nextB = c.nextA;
    }
    if (c.nextA >= 0) {
      Contact& nc = w.contacts[c.nextA >> 1];
      if ((c.nextA & 1) == 0) nc.prevA = c.prevA; else nc.prevB = c.prevA;
    }
    BodySim& ba = w.sims[c.bodyA];
    if (ba.headContactKey == ((id << 1) | 0)) ba.headContactKey = c.nextA;
    ba.contactCount -= 1;
  }
  {
    if (c.prevB >= 0) {
      Contact& pc = w.contacts[c.prevB >> 1];
      if ((c.prevB & 1) == 0) pc.nextA = c.nextB; else pc.nextB = c.nextB;
    }
    if (c.nextB >= 0) {
      Contact& nc = w.contacts[c.nextB >> 1];
      if ((c.nextB & 1) == 0) nc.prevA = c.prevB; else nc.prevB = c.prevB;
    }
    BodySim& bb = w.sims[c.bodyB];
    if (bb.headContactKey == ((id << 1) | 1)) bb.headContactKey = c.nextB;
    bb.contactCount -= 1;
  }
  if (c.color >= 0) removeContactFromGraph(w, id);
  c.shapeA = -1;
  c.shapeB = -1;
  c.flags = 0;
  c.simFlags = 0;
  w.links[id] = make_int4(-1, -1, -1, -1);
  WorldHeader* h = w.hdr;
  w.freeIds[h->freeCount++] = id;
}

// b2AssignJointColor + b2CreateJointInGraph, S/constraint_graph.c:214-290. The
// joint's index in its colour is the count of joints already there.
__device__ __noinline__ void addJointToGraph(const WorldPtrs& w, int jid) {
  Joint& j = w.joints[jid];
  const bool staticA = w.sims[j.bodyA].setIndex == kStaticSet;
  const bool staticB = w.sims[j.bodyB].setIndex == kStaticSet;
  int color = kOverflowIndex;
  for (int i = 0; i < kOverflowIndex; ++i) {
    uint32_t* set = colorSet(w, i);
    if (!staticA && !staticB) {
      if (getBit(set, j.bodyA) || getBit(set, j.bodyB)) continue;
      setBit(set, j.bodyA);
      setBit(set, j.bodyB);
    } else if (!staticA) {
      if (getBit(set, j.bodyA)) continue;
      setBit(set, j.bodyA);
    } else {
      if (getBit(set, j.bodyB)) continue;
      setBit(set, j.bodyB);
    }
    color = i;
    break;
  }
  int local = 0;
  for (int k = 0; k < w.hdr->jointCount; ++k) local += (k != jid && w.joints[k].color == color) ? 1 : 0;
  j.color = color;
  j.localIndex = local;
}

// b2RemoveJointFromGraph, S/constraint_graph.c:292-322: the colour's last joint
// takes the removed joint's index
__device__ __noinline__ void removeJointFromGraph(const WorldPtrs& w, int jid) {
  Joint& j = w.joints[jid];
  const int color = j.color;
  if (color < 0) return;
  if (color != kOverflowIndex) {
    uint32_t* set = colorSet(w, color);
    clearBit(set, j.bodyA);
    clearBit(set, j.bodyB);
  }
  int count = 0;
  for (int k = 0; k < w.hdr->jointCount; ++k) count += w.joints[k].color == color ? 1 : 0;
  const int last = count - 1;
  if (j.localIndex != last) {
    for (int k = 0; k < w.hdr->jointCount; ++k) {
      if (k != jid && w.joints[k].color == color && w.joints[k].localIndex == last) {
        w.joints[k].localIndex = j.localIndex;
        break;
      }
    }
  }
  j.color = -1;
  j.localIndex = -1;
}

// b2WakeSolverSet, S/solver_set.c:31-154. A sleeping set holds exactly one island
// (sets merge only through joints created between sleeping bodies); its bodies,
// touching contacts and joints are that island's lists in order.
__device__ __noinline__ void wakeIslandSet(const WorldPtrs& w, int islandId) {
  IslandCtx c = w.isl;
  WorldHeader* h = w.hdr;
  Island& is = c.islands[islandId];
  const int setIndex = is.setIndex;
  for (int b = is.headBody, g = h->bodyCount; b != -1 && g > 0; b = w.sims[b].islandNext, --g) {
    BodySim& body = w.sims[b];
    body.setIndex = kAwakeSet;
    body.localIndex = h->awakeBodyCount;
    c.awakeBodies[h->awakeBodyCount++] = b;
    body.sleepTime = 0.0f;
    w.states[b] = identityState();
    // its non-touching contacts leave the disabled set
    for (int key = body.headContactKey, gk = 2 * h->contactCap; key != -1 && gk > 0; --gk) {
      int cid = key >> 1;
      Contact& ct = w.contacts[cid];
      key = (key & 1) ? ct.nextB : ct.nextA;
      if (w.links[cid].w != kDisabledSet) continue;
      w.links[cid].w = kAwakeSet;
      ct.flags &= ~CF_OFF_AWAKE;
      ct.localIndex = -1;
    }
  }
  for (int q = is.headContact, g = h->contactCap; q != -1 && g > 0; q = w.links[q].z, --g) {
    w.contacts[q].localIndex = -1;
    addContactToGraph(w, q);
    w.links[q].w = kAwakeSet;
    w.contacts[q].flags &= ~CF_OFF_AWAKE;
  }
  for (int j = is.headJoint, g = h->jointCount; j != -1 && g > 0; j = w.joints[j].islandNext, --g) {
    addJointToGraph(w, j);
    w.joints[j].setIndex = kAwakeSet;
  }
  is.setIndex = kAwakeSet;
  is.localIndex = h->awakeIslandCount;
  c.awakeIslands[h->awakeIslandCount++] = islandId;
  c.setFree[h->setFreeCount++] = setIndex;  // b2DestroySolverSet
}

// b2LinkContact, S/island.c:117-218: a touching contact between an awake and a
// sleeping body wakes the sleeping set first
__device__ __noinline__ void linkContact(const WorldPtrs& w, int id) {
  const Contact& ct = w.contacts[id];
  const BodySim& A = w.sims[ct.bodyA];
  const BodySim& B = w.sims[ct.bodyB];
  if (A.setIndex == kAwakeSet && B.setIndex >= kFirstSleepingSet) wakeIslandSet(w, B.islandId);
  if (B.setIndex == kAwakeSet && A.setIndex >= kFirstSleepingSet) wakeIslandSet(w, A.islandId);
  IslandCtx c = w.isl;
  linkContactIslands(c, id);
}

// b2TrySleepIsland, S/solver_set.c:156-420: the island's bodies leave the awake
// set (remove-swap), its non-touching contacts to sleeping or static bodies go to
// the disabled set, its touching contacts and joints leave the constraint graph.
__device__ __noinline__ void trySleepIsland(const WorldPtrs& w, int islandId) {
  IslandCtx c = w.isl;
  WorldHeader* h = w.hdr;
  Island& is = c.islands[islandId];
  if (is.removeCount > 0) return;
  const int setId = poolAlloc(c.setFree, h->setFreeCount, h->setCount);
  int sleepIndex = 0;
  for (int b = is.headBody, g = h->bodyCount; b != -1 && g > 0; b = w.sims[b].islandNext, --g) {
    BodySim& body = w.sims[b];
    body.flags |= BF_FELL_ASLEEP;
    const int li = body.localIndex, last = --h->awakeBodyCount;
    if (li != last) {
      const int moved = c.awakeBodies[last];
      c.awakeBodies[li] = moved;
      w.sims[moved].localIndex = li;
    }
    body.setIndex = setId;
    body.localIndex = sleepIndex++;
    w.states[b] = identityState();
    for (int key = body.headContactKey, gk = 2 * h->contactCap; key != -1 && gk > 0; --gk) {
      const int cid = key >> 1, edge = key & 1;
      Contact& ct = w.contacts[cid];
      key = edge ? ct.nextB : ct.nextA;
      if (w.links[cid].w == kDisabledSet) continue;
      if (ct.color != -1) continue;
      const int other = edge ? ct.bodyA : ct.bodyB;
      if (w.sims[other].setIndex == kAwakeSet) continue;
      w.links[cid].w = kDisabledSet;
      ct.flags |= CF_OFF_AWAKE;
      ct.localIndex = -1;
    }
  }
  int k = 0;
  for (int q = is.headContact, g = h->contactCap; q != -1 && g > 0; q = w.links[q].z, --g) {
    removeContactFromGraph(w, q);
    w.links[q].w = setId;
    w.contacts[q].flags |= CF_OFF_AWAKE;
    w.contacts[q].localIndex = k++;
  }
  k = 0;
  for (int j = is.headJoint, g = h->jointCount; j != -1 && g > 0; j = w.joints[j].islandNext, --g) {
    removeJointFromGraph(w, j);
    w.joints[j].setIndex = setId;
    w.joints[j].localIndex = k++;
  }
  const int li = is.localIndex, last = --h->awakeIslandCount;
  if (li != last) {
    const int moved = c.awakeIslands[last];
    c.awakeIslands[li] = moved;
    c.islands[moved].localIndex = li;
  }
  is.setIndex = setId;
  is.localIndex = 0;
}

// Island sleep after finalize and continuous collision, S/solver.c:1995-2049: the
// split candidate for the next step, then every awake island no body kept awake,
// in reverse awake order (thread 0)
__device__ __noinline__ void sleepPhase(const WorldPtrs& w) {
  IslandCtx c = w.isl;
  WorldHeader* h = w.hdr;
  h->sleepPending = 0;
  int candidate = -1;
  const unsigned long long key = h->splitKey;
  if (key != 0ull) {
    const int simIndex = (int)(0xFFFFFFFFu - (uint32_t)key);
    if (simIndex >= 0 && simIndex < h->moveBodyCount) candidate = w.sims[c.moveOrder[simIndex]].islandId;
  }
  h->splitIslandId = candidate;
  for (int i = h->awakeIslandCount - 1; i >= 0; --i) {
    if ((c.awakeBits[i >> 5] >> (i & 31)) & 1u) continue;
    trySleepIsland(w, c.awakeIslands[i]);
  }
}

// Serial contact state loop, S/world.c:587-690
// What the serial state loop reads of one changed contact, gathered by all lanes
// before the loop. None of it changes during the loop except the bodies' solver
// sets, which a woken sleeping set moves (then they are read again).
struct ChangeRec {
  int id, flags, simFlags, shapeA, shapeB, bodyA, bodyB, islandA, islandB, setA, setB, bits;
};
enum : int { CR_STATIC_A = 1, CR_STATIC_B = 2, CR_LINKED = 4 };

// Contact state changes in ascending contact id (S/world.c:553-628): begin and end
// events, island links, graph colours, destroyed contacts. The changed contacts go
// through LDS in chunks: all lanes gather a chunk's records, then one lane applies
// them in order. serialW: a workgroup-shared copy of w for the out-of-line island
// calls, so the world pointers stay in registers elsewhere instead of a per-lane
// stack copy.
template <int BLOCK>
__device__ void statePhase(const WorldPtrs& w, SharedState& sh, WorldPtrs& serialW, char* lds, int ldsBytes) {
  const int n = sh.changedCount;
  const int* list = w.scratch;
  WorldHeader* h = w.hdr;
  ChangeRec* recs = reinterpret_cast<ChangeRec*>(lds);
  const int cap = max(ldsBytes / (int)sizeof(ChangeRec), 1);
  if (threadIdx.x == 0) serialW = w;
  int nBegin = 0, nEnd = 0;  // thread 0's
  for (int base = 0; base < n; base += cap) {
    const int m = min(cap, n - base);
    for (int k = threadIdx.x; k < m; k += BLOCK) {
      const int id = list[base + k];
      const Contact& c = w.contacts[id];
      ChangeRec r;
      r.id = id;
      r.flags = c.flags;
      r.simFlags = c.simFlags;
      r.shapeA = c.shapeA;
      r.shapeB = c.shapeB;
      r.bodyA = c.bodyA;
      r.bodyB = c.bodyB;
      const BodySim& A = w.sims[r.bodyA];
      const BodySim& B = w.sims[r.bodyB];
      r.islandA = A.islandId;
      r.islandB = B.islandId;
      r.setA = A.setIndex;
      r.setB = B.setIndex;
      r.bits = (A.type == kStatic ? CR_STATIC_A : 0) | (B.type == kStatic ? CR_STATIC_B : 0) |
               (w.links[id].x != -1 ? CR_LINKED : 0);
      recs[k] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int* beginIds = w.events;
      int* endPairs = w.events + h->contactCap;
      bool woke = false;
      for (int k = 0; k < m; ++k) {
        const ChangeRec r = recs[k];
        const int id = r.id;
        Contact& c = w.contacts[id];
        if (r.simFlags & SIM_DISJOINT) {
          // b2DestroyContact reports the end of a touching contact, S/contact.c:354-372
          if ((r.flags & CF_TOUCHING) && (r.flags & CF_CONTACT_EVENTS) && nEnd < h->contactCap) {
            endPairs[2 * nEnd] = r.shapeA;
            endPairs[2 * nEnd + 1] = r.shapeB;
            ++nEnd;
          }
          if (r.bits & CR_LINKED) {
            IslandCtx ic = w.isl;
            unlinkContact(ic, id);
          }
          destroyContact(w, id);
        } else if (r.simFlags & SIM_STARTED) {
          if ((r.flags & CF_CONTACT_EVENTS) && nBegin < h->contactCap) beginIds[nBegin++] = id;
          c.flags = r.flags | CF_TOUCHING;
          // b2LinkContact (S/island.c:117-218): a sleeping set touched from the awake
          // set wakes first
          int setA = r.setA, setB = r.setB;
          if (woke) {
            setA = w.sims[r.bodyA].setIndex;
            setB = w.sims[r.bodyB].setIndex;
          }
          if ((setA == kAwakeSet && setB >= kFirstSleepingSet) || (setB == kAwakeSet && setA >= kFirstSleepingSet)) {
            linkContact(serialW, id);
            woke = true;
          } else {
            IslandCtx ic = w.isl;
            linkContactIslandIds(ic, id, r.islandA, r.islandB);
          }
          c.simFlags = r.simFlags & ~SIM_STARTED;
          addContactToGraphBodies(w, id, r.bodyA, r.bodyB, (r.bits & CR_STATIC_A) != 0, (r.bits & CR_STATIC_B) != 0);
        } else if (r.simFlags & SIM_STOPPED) {
          c.simFlags = r.simFlags & ~SIM_STOPPED;
          c.flags = r.flags & ~CF_TOUCHING;
          if ((r.flags & CF_CONTACT_EVENTS) && nEnd < h->contactCap) {
            endPairs[2 * nEnd] = r.shapeA;
            endPairs[2 * nEnd + 1] = r.shapeB;
            ++nEnd;
          }
          if (r.bits & CR_LINKED) {
            IslandCtx ic = w.isl;
            unlinkContact(ic, id);
          }
          removeContactFromGraph(w, id);
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    h->beginCount = nBegin;
    h->endCount = nEnd;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------- phase: solve + finalize

struct SolveCtx {
  float h, inv_h, dt, inv_dt;
  int subSteps;
  Softness contactSoftness, staticSoftness;
  float contactSpeed, maxContactPushSpeed, restitutionThreshold, maxLinearSpeed, maxAngularSpeed;
  float warmStartScale;
  Vec2 gravity;
  bool enableContinuous, enableSleep;
};

template <int BLOCK, bool kLds>
__device__ void solveAndFinalize(const WorldPtrs& w, SharedState& sh, BodyState* ldsStates, const SolveCtx& ctx,
                                 int phases, int ldsBytes) {
  WorldHeader* h = w.hdr;
  const int nb = h->bodyCount;
  const int nc = h->contactCount;
  BodyState* states = kLds ? ldsStates : w.states;
  BodyConst* ldsConsts = reinterpret_cast<BodyConst*>(ldsStates + nb);  // LDS only
  // the constraint SoA: the world's lanes in LDS behind the body terms when they fit
  // the workgroup's LDS, else the world's region of the batch array in HBM; the
  // same for the joints
  Cons cons = w.cons;
  Joint* jts = w.joints;
  // joint ids by colour (behind the colour table of the stage launches in a split solve)
  int* jointLanes = (phases & PH_SOLVE_PREP) ? w.scratch + kWideTableInts : w.scratch;

  // b2Solve, S/solver.c:1190-1220 and :1479-1706: merge islands linked since the last
  // step, stop when no body is awake, split the last step's candidate island
  if (phases & PH_SOLVE) {
    // b2MergeAwakeIslands (S/island.c:547-590). Its first loop leaves every awake
    // island pointing at its root; here every lane walks its islands' parent chains
    // read-only (a chain read while other lanes shorten it still ends at the same
    // root) and writes its own island's root, then one lane merges in reverse
    // awake order.
    const bool merge = h->mergePending != 0;
    if (merge) {
      IslandCtx ic = w.isl;
      for (int i = threadIdx.x; i < h->awakeIslandCount; i += BLOCK) {
        const int id = ic.awakeIslands[i];
        int root = id;
        for (int guard = h->islandCount; ic.islands[root].parent != -1 && guard > 0; --guard)
          root = ic.islands[root].parent;
        if (root != id) ic.islands[id].parent = root;
      }
      __syncthreads();
      // the relabelling of b2MergeIsland on all lanes: whatever belongs to an awake
      // island with a parent (a child to merge) takes the parent's (its root's) id
      const int ni = h->islandCap;
      for (int i = threadIdx.x; i < nb; i += BLOCK) {
        const int id = w.sims[i].islandId;
        if (id < 0 || id >= ni) continue;
        const Island& is = ic.islands[id];
        if (is.parent != -1 && is.setIndex == kAwakeSet) w.sims[i].islandId = is.parent;
      }
      for (int k = threadIdx.x; k < nc; k += BLOCK) {
        const int id = w.links[k].x;
        if (id < 0 || id >= ni) continue;
        const Island& is = ic.islands[id];
        if (is.parent != -1 && is.setIndex == kAwakeSet) w.links[k].x = is.parent;
      }
      for (int j = threadIdx.x; j < h->jointCount; j += BLOCK) {
        const int id = w.joints[j].islandId;
        if (id < 0 || id >= ni) continue;
        const Island& is = ic.islands[id];
        if (is.parent != -1 && is.setIndex == kAwakeSet) w.joints[j].islandId = is.parent;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      IslandCtx ic = w.isl;
      if (merge) mergeAwakeChildren(ic);
      sh.noAwake = h->awakeBodyCount == 0 ? 1 : 0;
      if (!sh.noAwake && h->splitIslandId != -1) splitIsland(ic, h->splitIslandId);
      if (!sh.noAwake) h->splitIslandId = -1;
    }
    __syncthreads();
    if (sh.noAwake) return;
  } else if (phases & PH_FINALIZE) {
    if (threadIdx.x == 0) sh.noAwake = h->awakeBodyCount == 0 ? 1 : 0;
    __syncthreads();
    if (sh.noAwake) return;
  }

  if (phases & PH_SOLVE) {
    // -- bucket touching constraints by colour (overflow keeps its list order)
    if (threadIdx.x < kGraphColorCount) {
      sh.colorCount[threadIdx.x] = 0;
      sh.colorFill[threadIdx.x] = 0;
      sh.jointColorCount[threadIdx.x] = 0;
      sh.jointColorFill[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += BLOCK) {
      int color = w.ccolor[i];
      if (color >= 0 && color < kOverflowIndex) atomicAdd(&sh.colorCount[color], 1);
    }
    for (int i = threadIdx.x; i < h->jointCount; i += BLOCK) {
      const Joint& j = w.joints[i];
      if (j.color >= 0) atomicAdd(&sh.jointColorCount[j.color], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0, sj = 0, active = 0;
      for (int k = 0; k < kOverflowIndex; ++k) {
        sh.colorStart[k] = s;
        s += sh.colorCount[k];
        sh.jointColorStart[k] = sj;
        sj += sh.jointColorCount[k];
        active += (sh.colorCount[k] + sh.jointColorCount[k]) > 0 ? 1 : 0;
      }
      sh.colorStart[kOverflowIndex] = s;
      sh.colorCount[kOverflowIndex] = h->overflowCount;
      sh.colorStart[kGraphColorCount] = s + h->overflowCount;
      sh.jointColorStart[kOverflowIndex] = sj;
      h->activeColors = active;
    }
    __syncthreads();
    // LDS behind the body terms: the joints (copied in here, back after the
    // substeps), then the constraint lanes, as far as each fits
    if (kLds) {
      const long lanes = sh.colorStart[kGraphColorCount];
      long off = (long)nb * (long)(sizeof(BodyState) + sizeof(BodyConst));
      const long jointBytes = (long)h->jointCount * (long)sizeof(Joint);
      if (h->jointCount > 0 && off + jointBytes <= ldsBytes) {
        jts = reinterpret_cast<Joint*>(reinterpret_cast<char*>(ldsStates) + off);
        off += jointBytes;
      }
      if (lanes > 0 && off + 4L * CN_FIELD_COUNT * lanes <= ldsBytes) {
        cons.f = reinterpret_cast<float*>(reinterpret_cast<char*>(ldsStates) + off);
        cons.stride = lanes;
        off += 4L * CN_FIELD_COUNT * lanes;
      }
      // the joint ids by colour, read at every colour stage
      if (h->jointCount > 0 && off + 4L * h->jointCount <= ldsBytes) {
        jointLanes = reinterpret_cast<int*>(reinterpret_cast<char*>(ldsStates) + off);
      }
    }
    for (int i = threadIdx.x; i < nc; i += BLOCK) {
      int color = w.ccolor[i];
      if (color >= 0 && color < kOverflowIndex) {
        int lane = sh.colorStart[color] + atomicAdd(&sh.colorFill[color], 1);
        cons.seti(CN_SLOT, lane, i);
      }
    }
    for (int i = threadIdx.x; i < h->overflowCount; i += BLOCK) {
      cons.seti(CN_SLOT, sh.colorStart[kOverflowIndex] + i, w.overflow[i]);
    }
    for (int i = threadIdx.x; i < h->jointCount; i += BLOCK) {
      const Joint& j = w.joints[i];
      if (j.color >= 0 && j.color < kOverflowIndex) {
        jointLanes[sh.jointColorStart[j.color] + atomicAdd(&sh.jointColorFill[j.color], 1)] = i;
      }
    }
    if (threadIdx.x == 0 && sh.jointColorCount[kOverflowIndex] > 0) {
      // overflow joints are solved one after another in their colour's array order
      int* list = jointLanes + sh.jointColorStart[kOverflowIndex];
      const int n = sh.jointColorCount[kOverflowIndex];
      for (int k = 0; k < n; ++k) list[k] = -1;
      for (int i = 0; i < h->jointCount; ++i) {
        const Joint& j = w.joints[i];
        if (j.color == kOverflowIndex && j.localIndex >= 0 && j.localIndex < n) list[j.localIndex] = i;
      }
      for (int k = 0; k < n; ++k)
        if (list[k] < 0) h->error |= ERR_OVERFLOW_CAP;
    }
    // body states and per-step body terms into LDS
    if (kLds) {
      for (int i = threadIdx.x; i < nb; i += BLOCK) {
        ldsStates[i] = w.states[i];
        ldsConsts[i] = makeBodyConst(w.sims[i], ctx.gravity, ctx.h);
      }
      if (jts != w.joints) {
        const int words = h->jointCount * (int)(sizeof(Joint) / sizeof(float4));
        const float4* src = reinterpret_cast<const float4*>(w.joints);
        float4* dst = reinterpret_cast<float4*>(jts);
        for (int i = threadIdx.x; i < words; i += BLOCK) dst[i] = src[i];
      }
    }
    __syncthreads();
    if (phases & PH_SOLVE_PREP) {
      // the colour table of the stage launches (wideStageKernel)
      int* table = w.scratch;
      if (threadIdx.x < kGraphColorCount) {
        table[kWideColorCount + threadIdx.x] = sh.colorCount[threadIdx.x];
        table[kWideJointCount + threadIdx.x] = sh.jointColorCount[threadIdx.x];
        table[kWideJointStart + threadIdx.x] = sh.jointColorStart[threadIdx.x];
      }
      if (threadIdx.x <= kGraphColorCount) table[kWideColorStart + threadIdx.x] = sh.colorStart[threadIdx.x];
      return;
    }
    const int totalLanes = sh.colorStart[kGraphColorCount];
    const int overflowStart = sh.colorStart[kOverflowIndex];
    const int overflowCount = sh.colorCount[kOverflowIndex];
    const int coloredJoints = sh.jointColorStart[kOverflowIndex];
    const int overflowJoints = sh.jointColorCount[kOverflowIndex];
    const int* overflowJointList = jointLanes + coloredJoints;
    // the overflow stages (one lane, then a barrier) only when there is overflow
    const bool anyOverflow = overflowJoints + overflowCount > 0;

    // -- prepare joints (all), S/solver.c:967-987 -> S/joint.c:1348-1402
    for (int i = threadIdx.x; i < h->jointCount; i += BLOCK) {
      if (jts[i].color >= 0) prepareJoint(jts[i], w.sims, ctx.h, ctx.inv_h, ctx.warmStartScale);
    }
    // -- prepare contacts (all lanes), S/contact_solver.c:1438-1669
    for (int lane = threadIdx.x; lane < totalLanes; lane += BLOCK) {
      int slot = cons.geti(CN_SLOT, lane);
      const Contact& c = w.contacts[slot];
      bool staticA, staticB;
      float mA, iA, mB, iB;
      if (kLds) {
        const BodyConst& ka = ldsConsts[c.bodyA];
        const BodyConst& kb = ldsConsts[c.bodyB];
        staticA = (ka.flags & BC_STATIC) != 0;
        staticB = (kb.flags & BC_STATIC) != 0;
        mA = ka.invMass, iA = ka.invI, mB = kb.invMass, iB = kb.invI;
      } else {
        const BodySim& sa = w.sims[c.bodyA];
        const BodySim& sb = w.sims[c.bodyB];
        staticA = sa.type == kStatic;
        staticB = sb.type == kStatic;
        mA = sa.invMass, iA = sa.invI, mB = sb.invMass, iB = sb.invI;
      }
      int ia = staticA ? -1 : c.bodyA;
      int ib = staticB ? -1 : c.bodyB;
      if (staticA) mA = 0.0f, iA = 0.0f;
      if (staticB) mB = 0.0f, iB = 0.0f;
      prepareContactLane(cons, lane, c, slot, ia, ib, mA, iA, mB, iB, states, ctx.contactSoftness,
                         ctx.staticSoftness, ctx.warmStartScale);
    }
    __syncthreads();

    for (int sub = 0; sub < ctx.subSteps; ++sub) {
      // integrate velocities
      for (int i = threadIdx.x; i < nb; i += BLOCK) {
        if (kLds) {
          BodyConst k = ldsConsts[i];
          if (k.flags & BC_STATIC) continue;
          BodyState st = states[i];
          integrateVelocityConst(k, st, ctx.maxLinearSpeed, ctx.maxAngularSpeed);
          states[i] = st;
          ldsConsts[i].flags = k.flags;
        } else {
          BodySim& sim = w.sims[i];
          if (sim.setIndex != kAwakeSet) continue;
          BodyState st = states[i];
          integrateVelocity(sim, st, ctx.gravity, ctx.h, ctx.maxLinearSpeed, ctx.maxAngularSpeed);
          states[i] = st;
        }
      }
      __syncthreads();
      // warm start: overflow joints, overflow contacts, then colours
      if (anyOverflow) {
        if (threadIdx.x == 0) {
          for (int k = 0; k < overflowJoints; ++k)
            if (overflowJointList[k] >= 0) warmStartJoint(jts[overflowJointList[k]], states);
          for (int k = 0; k < overflowCount; ++k) overflowWarmStart(cons, overflowStart + k, states);
        }
        __syncthreads();
      }
      for (int col = 0; col < kOverflowIndex; ++col) {
        int cnt = sh.colorCount[col], jcnt = sh.jointColorCount[col];
        if (cnt + jcnt == 0) continue;
        for (int k = threadIdx.x; k < jcnt; k += BLOCK)
          warmStartJoint(jts[jointLanes[sh.jointColorStart[col] + k]], states);
        for (int k = threadIdx.x; k < cnt; k += BLOCK) warmStartLane(cons, sh.colorStart[col] + k, states);
        __syncthreads();
      }
      // solve
      if (anyOverflow) {
        if (threadIdx.x == 0) {
          for (int k = 0; k < overflowJoints; ++k)
            if (overflowJointList[k] >= 0)
              solveJoint(jts[overflowJointList[k]], states, ctx.h, ctx.inv_h, true);
          for (int k = 0; k < overflowCount; ++k)
            overflowSolve(cons, overflowStart + k, states, ctx.inv_h, ctx.maxContactPushSpeed, true);
        }
        __syncthreads();
      }
      for (int col = 0; col < kOverflowIndex; ++col) {
        int cnt = sh.colorCount[col], jcnt = sh.jointColorCount[col];
        if (cnt + jcnt == 0) continue;
        for (int k = threadIdx.x; k < jcnt; k += BLOCK)
          solveJoint(jts[jointLanes[sh.jointColorStart[col] + k]], states, ctx.h, ctx.inv_h, true);
        for (int k = threadIdx.x; k < cnt; k += BLOCK)
          solveLane(cons, sh.colorStart[col] + k, states, ctx.inv_h, -ctx.contactSpeed, true);
        __syncthreads();
      }
      // integrate positions
      for (int i = threadIdx.x; i < nb; i += BLOCK) {
        if (kLds ? (ldsConsts[i].flags & BC_STATIC) != 0 : w.sims[i].setIndex != kAwakeSet) continue;
        BodyState st = states[i];
        integratePosition(st, ctx.h);
        states[i] = st;
      }
      __syncthreads();
      // relax
      if (anyOverflow) {
        if (threadIdx.x == 0) {
          for (int k = 0; k < overflowJoints; ++k)
            if (overflowJointList[k] >= 0)
              solveJoint(jts[overflowJointList[k]], states, ctx.h, ctx.inv_h, false);
          for (int k = 0; k < overflowCount; ++k)
            overflowSolve(cons, overflowStart + k, states, ctx.inv_h, ctx.maxContactPushSpeed, false);
        }
        __syncthreads();
      }
      for (int col = 0; col < kOverflowIndex; ++col) {
        int cnt = sh.colorCount[col], jcnt = sh.jointColorCount[col];
        if (cnt + jcnt == 0) continue;
        for (int k = threadIdx.x; k < jcnt; k += BLOCK)
          solveJoint(jts[jointLanes[sh.jointColorStart[col] + k]], states, ctx.h, ctx.inv_h, false);
        for (int k = threadIdx.x; k < cnt; k += BLOCK)
          solveLane(cons, sh.colorStart[col] + k, states, ctx.inv_h, -ctx.contactSpeed, false);
        __syncthreads();
      }
    }
    // the joints' impulses and prepared data back to HBM
    if (jts != w.joints) {
      const int words = h->jointCount * (int)(sizeof(Joint) / sizeof(float4));
      const float4* src = reinterpret_cast<const float4*>(jts);
      float4* dst = reinterpret_cast<float4*>(w.joints);
      for (int i = threadIdx.x; i < words; i += BLOCK) dst[i] = src[i];
    }
    // restitution: overflow then colours
    if (overflowCount > 0) {
      if (threadIdx.x == 0) {
        for (int k = 0; k < overflowCount; ++k)
          overflowRestitution(cons, overflowStart + k, states, ctx.restitutionThreshold);
      }
      __syncthreads();
    }
    for (int col = 0; col < kOverflowIndex; ++col) {
      int cnt = sh.colorCount[col];
      if (cnt == 0) continue;
      for (int k = threadIdx.x; k < cnt; k += BLOCK)
        restitutionLane(cons, sh.colorStart[col] + k, states, ctx.restitutionThreshold);
      __syncthreads();
    }
    // store impulses
    for (int lane = threadIdx.x; lane < totalLanes; lane += BLOCK) {
      int slot = cons.geti(CN_SLOT, lane);
      if (lane >= overflowStart) {
        overflowStore(cons, lane, w.contacts[slot]);
      } else {
        storeLane(cons, lane, w.contacts[slot]);
      }
    }
    (void)coloredJoints;
    __syncthreads();
  } else if (kLds) {
    for (int i = threadIdx.x; i < nb; i += BLOCK) ldsStates[i] = w.states[i];
    __syncthreads();
  }

  if (phases & PH_FINALIZE) {
    // b2FinalizeBodiesTask, S/solver.c:543-726
    const IslandCtx& ic = w.isl;
    if (threadIdx.x == 0) {
      sh.fastCount = 0;
      h->splitKey = 0ull;
      h->moveBodyCount = h->awakeBodyCount;
    }
    for (int k = threadIdx.x; k < (h->awakeIslandCount + 31) / 32; k += BLOCK) ic.awakeBits[k] = 0u;
    // the awake order of this finalize (body move events follow it)
    for (int k = threadIdx.x; k < h->awakeBodyCount; k += BLOCK) ic.moveOrder[k] = ic.awakeBodies[k];
    __syncthreads();
    int* fastList = w.scratch;  // fast bodies for the continuous-collision launch
    for (int i = threadIdx.x; i < nb; i += BLOCK) {
      BodySim& sim = w.sims[i];
      if (sim.setIndex != kAwakeSet) continue;
      BodyState st = states[i];
      Vec2 v = st.v;
      float wv = st.w;
      sim.center = add(sim.center, st.dp);
      sim.xf.q = normalizeRot(mulRot(st.dq, sim.xf.q));
      float maxVelocity = length(v) + absf(wv) * sim.maxExtent;
      float maxDeltaPosition = length(st.dp) + absf(st.dq.s) * sim.maxExtent;
      float positionSleepFactor = 0.5f;
      float sleepVelocity = maxf(maxVelocity, positionSleepFactor * ctx.inv_dt * maxDeltaPosition);
      st.dp = {0.0f, 0.0f};
      st.dq = {1.0f, 0.0f};
      sim.xf.p = sub(sim.center, rotate(sim.xf.q, sim.localCenter));
      sim.force = {0.0f, 0.0f};
      sim.torque = 0.0f;
      // isSpeedCapped reports this step's capping (body->isSpeedCapped = sim->isSpeedCapped)
      bool capped = kLds ? (ldsConsts[i].flags & BC_CAPPED) != 0 : (sim.flags & BF_CAPPED_STEP) != 0;
      int flags = sim.flags & ~(BF_IS_FAST | BF_ENLARGED | BF_SPEED_CAPPED | BF_CAPPED_STEP | BF_FELL_ASLEEP);
      if (capped) flags |= BF_SPEED_CAPPED;
      bool isFast = false;
      if (!ctx.enableSleep || (sim.flags & BF_ENABLE_SLEEP) == 0 || sleepVelocity > sim.sleepThreshold) {
        sim.sleepTime = 0.0f;
        if (sim.type == kDynamic && ctx.enableContinuous && maxVelocity * ctx.dt > 0.5f * sim.minExtent) {
          // the continuous-collision launch sweeps this body (b2SolveContinuous)
          isFast = true;
          int slot = atomicAdd(&sh.fastCount, 1);
          fastList[slot] = i;
        } else {
          sim.center0 = sim.center;
          sim.rotation0 = sim.xf.q;
        }
      } else {
        sim.center0 = sim.center;
        sim.rotation0 = sim.xf.q;
        sim.sleepTime += ctx.dt;
      }
      // any body of an island keeps it awake; the sleepiest body (first in the awake
      // order on ties) of an island with removed constraints names the island to
      // split next step (S/solver.c:658-676)
      if (ctx.enableSleep && sim.islandId >= 0 && sim.islandId < h->islandCap) {
        const Island& is = ic.islands[sim.islandId];
        if (sim.sleepTime < kTimeToSleep) {
          atomicOr(&ic.awakeBits[is.localIndex >> 5], 1u << (is.localIndex & 31));
        } else if (is.removeCount > 0) {
          unsigned long long key = ((unsigned long long)__float_as_uint(sim.sleepTime) << 32) |
                                   (unsigned long long)(0xFFFFFFFFu - (uint32_t)sim.localIndex);
          atomicMax(&h->splitKey, key);
        }
      }
      Xf xf = sim.xf;
      int rank = 0;
      bool enlargedBody = false;
      const bool bullet = (sim.flags & BF_BULLET) != 0;
      for (int s = sim.headShape; s >= 0 && rank < h->shapeCount; s = w.shapes[s].next, ++rank) {
        Shape& shape = w.shapes[s];
        if (isFast) {
          // fast shapes get their AABBs from the continuous sweep; a fast bullet
          // joins the move buffer with every shape (S/solver.c:1866-1879)
          if (bullet) {
            shape.flags |= SF_MOVED;
            shape.moveRank = (sim.localIndex << 8) | (rank < 255 ? rank : 255);
            enlargedBody = true;
          }
          continue;
        }
        AABB aabb = computeShapeAABB(w.geoms[s], xf);
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
          shape.flags |= SF_TREE_PENDING;
          // enlarged proxy joins the move buffer in sim order (S/solver.c:1848-1903)
          shape.flags |= SF_MOVED;
          shape.moveRank = (sim.localIndex << 8) | (rank < 255 ? rank : 255);
          enlargedBody = true;
        }
      }
      if (enlargedBody) flags |= BF_ENLARGED;
      if (isFast) flags |= BF_IS_FAST;
      sim.flags = flags;
      w.states[i] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      h->fastCount = sh.fastCount;
      h->sleepPending = ctx.enableSleep ? 1 : 0;
    }
  } else if (kLds && (phases & PH_SOLVE)) {
    for (int i = threadIdx.x; i < nb; i += BLOCK) {
      w.states[i] = ldsStates[i];
      if (ldsConsts[i].flags & BC_CAPPED) w.sims[i].flags |= BF_CAPPED_STEP;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------- kernels

__device__ void stepContext(const WorldHeader* h, const StepParams& sp, SolveCtx& ctx) {
  ctx.dt = sp.dt;
  ctx.inv_dt = sp.inv_dt;
  ctx.h = sp.h;
  ctx.inv_h = sp.inv_h;
  ctx.subSteps = sp.subSteps;
  // S/world.c:760-771
  float contactHertz = minf(h->contactHertz, 0.125f * sp.inv_h);
  ctx.contactSoftness = makeSoft(contactHertz, h->contactDampingRatio, sp.h);
  ctx.staticSoftness = makeSoft(2.0f * contactHertz, h->contactDampingRatio, sp.h);
  ctx.contactSpeed = h->maxContactPushSpeed / ctx.staticSoftness.massScale;
  ctx.maxContactPushSpeed = h->maxContactPushSpeed;
  ctx.restitutionThreshold = h->restitutionThreshold;
  ctx.maxLinearSpeed = h->maxLinearSpeed;
  ctx.maxAngularSpeed = kMaxRotation * sp.inv_dt;
  ctx.warmStartScale = (h->flags & WF_WARM) ? 1.0f : 0.0f;
  ctx.gravity = h->gravity;
  ctx.enableContinuous = (h->flags & WF_CONTINUOUS) != 0;
  ctx.enableSleep = (h->flags & WF_SLEEP) != 0;
}

// Pairs, collide and contact state of one world per workgroup. ldsBytes of dynamic
// LDS hold the sort-and-sweep arrays (28 B per shape) with the existing-pair hash
// set behind them, then the new-pair sort (12 B per pair); larger worlds use their
// global scratch region and walk the contact lists instead.
// Four waves per SIMD: the compiler keeps the manifold code within 128 VGPRs
// (measured faster than the 148-VGPR, three-wave build on the 16k-world workload).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void collideKernel(BatchView b, StepParams sp, const int* worldList, int ldsBytes) {
  __shared__ SharedState sh;
  __shared__ WorldPtrs serialW;
  extern __shared__ __align__(16) char lds[];
  const int wi = worldList ? worldList[blockIdx.x] : blockIdx.x;
  WorldHeader* h = b.worlds + wi;
  WorldPtrs w = worldPtrs(b, h);
  if (threadIdx.x == 0) {
    SolveCtx ctx;
    stepContext(h, sp, ctx);
    sh.error = 0;
    h->inv_h = sp.inv_h;
    h->contactSpeed = ctx.contactSpeed;
    h->moveBodyCount = 0;  // body move events start empty each step (S/world.c:709)
  }
  __syncthreads();
  int ldsShapes = 1;
  while (ldsShapes * 2 * 28 <= ldsBytes) ldsShapes *= 2;
  if (ldsShapes * 28 > ldsBytes) ldsShapes = 0;
  int ldsPairs = 1;
  while (ldsPairs * 2 * 12 <= ldsBytes) ldsPairs *= 2;
  if (ldsPairs * 12 > ldsBytes) ldsPairs = 0;
  if (sp.phases & PH_PAIRS) {
    pairsPhase<BLOCK>(w, sh, lds, ldsShapes, ldsPairs, ldsBytes, (sp.phases & PH_PAIRS_PREP) ? 1 : 0);
  } else if (sp.phases & PH_PAIRS_FINISH) {
    pairsPhase<BLOCK>(w, sh, lds, ldsShapes, ldsPairs, ldsBytes, 2);
  }
  if (sp.phases & PH_COLLIDE) collidePhase<BLOCK>(w, sh, (sp.phases & PH_CONTACTS_DONE) != 0);
  if (sp.phases & PH_STATE) statePhase<BLOCK>(w, sh, serialW, lds, ldsBytes);
  if (threadIdx.x == 0 && (sp.phases & (PH_SOLVE | PH_FINALIZE | PH_NO_COUNT)) == 0) h->stepIndex += 1;
}

// The sweep of the worlds a mode-1 pair update left in HBM (sweepCount > 0): a
// workgroup per 256 sorted shapes (blockIdx.x), blockIdx.y the world of the class;
// new pairs go to the world's pair buffer through a global counter.
__global__ __launch_bounds__(256) void sweepKernel(BatchView b, const int* worldList) {
  __shared__ LongSweeps sh;
  const int wi = worldList ? worldList[blockIdx.y] : blockIdx.y;
  WorldHeader* h = b.worlds + wi;
  const int n = h->sweepCount;
  const int first = blockIdx.x * 256;
  if (first >= n) return;  // the whole workgroup: n is uniform
  WorldPtrs w = worldPtrs(b, h);
  const int np = nextPow2(max(n, 4));
  sweepShapes<256>(w, sh, w.sapBoxes, w.scratch + 2 * np, nullptr, first, min(256, n - first), n, &h->sweepPairs,
                   nullptr, 0u);
}

// Solver stages and finalize of one world per workgroup; body states and the
// per-step body terms live in ldsBytes of dynamic LDS (64 B per body) when they fit.
// The narrow phase of the worlds of a split pair update: workgroup blockIdx.x of
// gridDim.x takes every gridDim.x-th run of 256 contact slots of world blockIdx.y
// (contacts update independently; the collide launch after it lists the changes).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void narrowKernel(BatchView b,
                                                                                            const int* worldList) {
  const int wi = worldList ? worldList[blockIdx.y] : blockIdx.y;
  WorldHeader* h = b.worlds + wi;
  const int n = h->contactCount;
  if ((int)blockIdx.x * 256 >= n) return;
  WorldPtrs w = worldPtrs(b, h);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) collideContact(w, i);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void solveKernel(BatchView b, StepParams sp, const int* worldList, int ldsBytes) {
  __shared__ SharedState sh;
  extern __shared__ __align__(16) char lds[];
  const int wi = worldList ? worldList[blockIdx.x] : blockIdx.x;
  WorldHeader* h = b.worlds + wi;
  WorldPtrs w = worldPtrs(b, h);
  SolveCtx ctx;
  stepContext(h, sp, ctx);
  if (h->bodyCount * (int)(sizeof(BodyState) + sizeof(BodyConst)) <= ldsBytes) {
    solveAndFinalize<BLOCK, true>(w, sh, reinterpret_cast<BodyState*>(lds), ctx, sp.phases, ldsBytes);
  } else {
    solveAndFinalize<BLOCK, false>(w, sh, nullptr, ctx, sp.phases, ldsBytes);
  }
  if (threadIdx.x == 0 && (sp.phases & PH_NO_COUNT) == 0) h->stepIndex += 1;
}

// The solve stages of large worlds, one launch per stage (and colour), workgroup
// blockIdx.x of gridDim.x taking every gridDim.x-th run of 256 items of world
// blockIdx.y: the per-item work of solveAndFinalize's HBM path, in the same stage
// order (prepare; per substep: integrate velocities, warm start, solve, integrate
// positions, relax, each constraint stage the overflow first, one lane, then the
// colours; restitution; store). The bodies a colour's constraints touch are
// distinct, so a colour's items run in any order. After a PH_SOLVE_PREP launch.
enum WideStage : int { WS_PREPARE, WS_INTEGRATE_VEL, WS_WARM, WS_SOLVE, WS_INTEGRATE_POS, WS_RELAX, WS_RESTITUTION,
                       WS_STORE };

__global__ __launch_bounds__(256) void wideStageKernel(BatchView b, StepParams sp, const int* worldList, int stage,
                                                       int color) {
  const int wi = worldList ? worldList[blockIdx.y] : blockIdx.y;
  WorldHeader* h = b.worlds + wi;
  if (h->awakeBodyCount == 0) return;  // b2Solve stopped after merging islands
  const int* table = b.scratch + h->scratchOffset;
  const int gtid = blockIdx.x * 256 + threadIdx.x, stride = gridDim.x * 256;
  const bool constraintStage = stage == WS_WARM || stage == WS_SOLVE || stage == WS_RELAX || stage == WS_RESTITUTION;
  if (constraintStage) {
    const int cnt = table[kWideColorCount + color];
    const int jcnt = stage == WS_RESTITUTION ? 0 : table[kWideJointCount + color];
    if (color == kOverflowIndex ? (gtid != 0 || cnt + jcnt == 0) : gtid >= max(cnt, jcnt)) return;
  }
  WorldPtrs w = worldPtrs(b, h);
  SolveCtx ctx;
  stepContext(h, sp, ctx);
  const Cons cons = w.cons;
  BodyState* states = w.states;
  Joint* jts = w.joints;
  const int* jointLanes = w.scratch + kWideTableInts;
  const int nb = h->bodyCount;
  if (stage == WS_PREPARE) {
    // S/solver.c:967-987, S/contact_solver.c:1438-1669
    for (int i = gtid; i < h->jointCount; i += stride)
      if (jts[i].color >= 0) prepareJoint(jts[i], w.sims, ctx.h, ctx.inv_h, ctx.warmStartScale);
    const int totalLanes = table[kWideColorStart + kGraphColorCount];
    for (int lane = gtid; lane < totalLanes; lane += stride) {
      const int slot = cons.geti(CN_SLOT, lane);
      const Contact& c = w.contacts[slot];
      const BodySim& sa = w.sims[c.bodyA];
      const BodySim& sb = w.sims[c.bodyB];
      const bool staticA = sa.type == kStatic, staticB = sb.type == kStatic;
      float mA = sa.invMass, iA = sa.invI, mB = sb.invMass, iB = sb.invI;
      const int ia = staticA ? -1 : c.bodyA;
      const int ib = staticB ? -1 : c.bodyB;
      if (staticA) mA = 0.0f, iA = 0.0f;
      if (staticB) mB = 0.0f, iB = 0.0f;
      prepareContactLane(cons, lane, c, slot, ia, ib, mA, iA, mB, iB, states, ctx.contactSoftness,
                         ctx.staticSoftness, ctx.warmStartScale);
    }
  } else if (stage == WS_INTEGRATE_VEL) {
    for (int i = gtid; i < nb; i += stride) {
      BodySim& sim = w.sims[i];
      if (sim.setIndex != kAwakeSet) continue;
      BodyState st = states[i];
      integrateVelocity(sim, st, ctx.gravity, ctx.h, ctx.maxLinearSpeed, ctx.maxAngularSpeed);
      states[i] = st;
    }
  } else if (stage == WS_INTEGRATE_POS) {
    for (int i = gtid; i < nb; i += stride) {
      if (w.sims[i].setIndex != kAwakeSet) continue;
      BodyState st = states[i];
      integratePosition(st, ctx.h);
      states[i] = st;
    }
  } else if (stage == WS_STORE) {
    const int totalLanes = table[kWideColorStart + kGraphColorCount];
    const int overflowStart = table[kWideColorStart + kOverflowIndex];
    for (int lane = gtid; lane < totalLanes; lane += stride) {
      const int slot = cons.geti(CN_SLOT, lane);
      if (lane >= overflowStart) {
        overflowStore(cons, lane, w.contacts[slot]);
      } else {
        storeLane(cons, lane, w.contacts[slot]);
      }
    }
  } else if (color == kOverflowIndex) {
    // one lane, the overflow joints then the overflow contacts in list order
    const int start = table[kWideColorStart + kOverflowIndex], cnt = table[kWideColorCount + kOverflowIndex];
    const int* jointList = jointLanes + table[kWideJointStart + kOverflowIndex];
    const int jcnt = table[kWideJointCount + kOverflowIndex];
    if (stage == WS_RESTITUTION) {
      for (int k = 0; k < cnt; ++k) overflowRestitution(cons, start + k, states, ctx.restitutionThreshold);
    } else if (stage == WS_WARM) {
      for (int k = 0; k < jcnt; ++k)
        if (jointList[k] >= 0) warmStartJoint(jts[jointList[k]], states);
      for (int k = 0; k < cnt; ++k) overflowWarmStart(cons, start + k, states);
    } else {
      const bool useBias = stage == WS_SOLVE;
      for (int k = 0; k < jcnt; ++k)
        if (jointList[k] >= 0) solveJoint(jts[jointList[k]], states, ctx.h, ctx.inv_h, useBias);
      for (int k = 0; k < cnt; ++k)
        overflowSolve(cons, start + k, states, ctx.inv_h, ctx.maxContactPushSpeed, useBias);
    }
  } else {
    const int start = table[kWideColorStart + color], cnt = table[kWideColorCount + color];
    const int jstart = table[kWideJointStart + color], jcnt = table[kWideJointCount + color];
    if (stage == WS_RESTITUTION) {
      for (int k = gtid; k < cnt; k += stride) restitutionLane(cons, start + k, states, ctx.restitutionThreshold);
    } else if (stage == WS_WARM) {
      for (int k = gtid; k < jcnt; k += stride) warmStartJoint(jts[jointLanes[jstart + k]], states);
      for (int k = gtid; k < cnt; k += stride) warmStartLane(cons, start + k, states);
    } else {
      const bool useBias = stage == WS_SOLVE;
      for (int k = gtid; k < jcnt; k += stride)
        solveJoint(jts[jointLanes[jstart + k]], states, ctx.h, ctx.inv_h, useBias);
      for (int k = gtid; k < cnt; k += stride)
        solveLane(cons, start + k, states, ctx.inv_h, -ctx.contactSpeed, useBias);
    }
  }
}

// ---------------------------------------------------------------------------- phase: continuous

// b2GetShapeCentroid, S/shape.c:601-618
__device__ Vec2 shapeCentroid(const ShapeGeom& g) {
  if (g.type == kCircle) return g.v[0];
  if (g.type == kPolygon) return g.centroid;
  return lerp(g.v[0], g.v[1], 0.5f);
}

// minExtent of b2ComputeShapeExtent, S/shape.c:722-790
__device__ float shapeMinExtent(const ShapeGeom& g) {
  if (g.type == kCircle || g.type == kCapsule) return g.radius;
  if (g.type == kSegment) return 0.0f;
  float mn = kHuge;
  for (int i = 0; i < g.count && i < kMaxPolygonVertices; ++i) mn = minf(mn, dot(g.n[i], sub(g.v[i], g.centroid)));
  return mn + g.radius;
}

// b2ContinuousQueryCallback, S/solver.c:212-388, for candidate shape t of a fast
// shape s on body i. Returns the (possibly reduced) sweep fraction. Not inlined:
// the time-of-impact proxies (dynamically indexed arrays) live in this call's frame,
// so only lanes that reach a candidate touch them.
__device__ __noinline__ float continuousCandidate(const Shape* shapes, const ShapeGeom* geoms, const BodySim* sims,
                                                  const Joint* joints, int i, int s, Sweep sweep, int t,
                                                  float fraction) {
  if (t == s) return fraction;
  const Shape& fs = shapes[s];
  const Shape& ts = shapes[t];
  if (ts.body == i) return fraction;
  if (ts.flags & SF_SENSOR) return fraction;
  if (!shouldShapesCollide(fs, ts)) return fraction;
  const BodySim& other = sims[ts.body];
  if (other.flags & BF_BULLET) return fraction;
  if (!shouldBodiesCollide(sims, joints, i, ts.body)) return fraction;
  Proxy proxyB = makeShapeProxy(geoms[s]);
  Proxy proxyA = makeShapeProxy(geoms[t]);
  Sweep sweepA = makeSweep(other);
  float f = timeOfImpact(proxyA, sweepA, proxyB, sweep, fraction);
  if (0.0f < f && f < fraction) return f;
  if (f == 0.0f) {
    // fallback: a small circle around the fast shape's centroid
    const ShapeGeom& g = geoms[s];
    Proxy core;
    core.count = 1;
    core.p[0] = shapeCentroid(g);
    core.radius = 0.25f * shapeMinExtent(g);
    f = timeOfImpact(proxyA, sweepA, core, sweep, fraction);
    if (0.0f < f && f < fraction) return f;
  }
  return fraction;
}

__device__ __forceinline__ bool queryVisible(const Shape& t, AABB box) {
  // b2DynamicTree_Query with B2_DEFAULT_MASK_BITS against the proxy's fat AABB
  return (t.catLo | t.catHi) != 0 && t.proxyKey >= 0 && aabbOverlaps(t.fat, box);
}

// b2SolveContinuous, S/solver.c:390-541, for fast body i. Non-bullets sweep against
// the static proxies; bullets also against kinematic and dynamic proxies, after
// every non-bullet body has its final transform. Candidates are visited in each
// tree's query order (b2DynamicTree_Query, S/dynamic_tree.c:1114-1170: the leaves
// whose fat boxes overlap, in depth-first order): each hit bounds the next
// candidate's root search, so the order fixes the fraction's bits when several hit.
// The static proxies' fat boxes and visibility staged in LDS by the continuous
// kernel (the first kMaxStagedStatics; the rest are read from their records)
constexpr int kMaxStagedStatics = 64;
constexpr int kContinuousLanes = 64;  // fast bodies of one world per workgroup pass
struct StaticProxies {
  AABB fat[kMaxStagedStatics];
  int visible[kMaxStagedStatics];
};

__device__ void solveContinuous(const WorldPtrs& w, int i, const int* staticList, int staticCount, bool bullet,
                                const StaticProxies& statics) {
  BodySim& sim = w.sims[i];
  const int ns = w.hdr->shapeCount;
  Sweep sweep = makeSweep(sim);
  Xf xf2 = {sub(sweep.c2, rotate(sweep.q2, sweep.localCenter)), sweep.q2};
  float fraction = 1.0f;
  int rank = 0;
  for (int s = sim.headShape; s >= 0 && rank < ns; s = w.shapes[s].next, ++rank) {
    Shape& fs = w.shapes[s];
    AABB box1 = fs.aabb;
    AABB box2 = computeShapeAABB(w.geoms[s], xf2);
    AABB box;
    box.lo = vmin(box1.lo, box2.lo);
    box.hi = vmax(box1.hi, box2.hi);
    fs.aabb = box2;
    if (fs.flags & SF_SENSOR) continue;
    for (int k = 0; k < staticCount; ++k) {
      const bool staged = k < kMaxStagedStatics;
      const int t = staticList[k];
      if (staged ? !(statics.visible[k] && aabbOverlaps(statics.fat[k], box)) : !queryVisible(w.shapes[t], box))
        continue;
      fraction = continuousCandidate(w.shapes, w.geoms, w.sims, w.joints, i, s, sweep, t, fraction);
    }
    if (bullet) {
      // the visible proxies of each tree by ascending query rank; a scan finds the
      // next one (bullets are few)
      for (int type = kKinematic; type <= kDynamic; ++type) {
        int last = -1;
        for (int n = 0; n < ns; ++n) {
          int next = -1, nextRank = 0x7FFFFFFF;
          for (int t = 0; t < ns; ++t) {
            const Shape& ts = w.shapes[t];
            if ((ts.proxyKey & 3) != type || !queryVisible(ts, box)) continue;
            const int r = w.treeRanks[t];
            if (r > last && r < nextRank) {
              nextRank = r;
              next = t;
            }
          }
          if (next < 0) break;
          fraction = continuousCandidate(w.shapes, w.geoms, w.sims, w.joints, i, s, sweep, next, fraction);
          last = nextRank;
        }
      }
    }
  }
  int flags = sim.flags;
  rank = 0;
  if (fraction < 1.0f) {
    // time of impact: rewind the body to the hit fraction
    Rot q = nlerpRot(sweep.q1, sweep.q2, fraction);
    Vec2 c = lerp(sweep.c1, sweep.c2, fraction);
    Vec2 origin = sub(c, rotate(q, sweep.localCenter));
    Xf xf = {origin, q};
    sim.xf = xf;
    sim.center = c;
    sim.rotation0 = q;
    sim.center0 = c;
    for (int s = sim.headShape; s >= 0 && rank < ns; s = w.shapes[s].next, ++rank) {
      Shape& shape = w.shapes[s];
      AABB aabb = computeShapeAABB(w.geoms[s], xf);
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
        shape.flags |= SF_TREE_PENDING;
        shape.flags |= SF_MOVED;
        shape.moveRank = (sim.localIndex << 8) | (rank < 255 ? rank : 255);
        flags |= BF_ENLARGED;
      }
    }
  } else {
    sim.rotation0 = sim.xf.q;
    sim.center0 = sim.center;
    for (int s = sim.headShape; s >= 0 && rank < ns; s = w.shapes[s].next, ++rank) {
      Shape& shape = w.shapes[s];
      // the shape AABB is the end-of-sweep box computed above (no speculative margin)
      if (!aabbContains(shape.fat, shape.aabb)) {
        AABB fat;
        fat.lo.x = shape.aabb.lo.x - kAabbMargin;
        fat.lo.y = shape.aabb.lo.y - kAabbMargin;
        fat.hi.x = shape.aabb.hi.x + kAabbMargin;
        fat.hi.y = shape.aabb.hi.y + kAabbMargin;
        shape.fat = fat;
        shape.flags |= SF_TREE_PENDING;
        shape.flags |= SF_MOVED;
        shape.moveRank = (sim.localIndex << 8) | (rank < 255 ? rank : 255);
        flags |= BF_ENLARGED;
      }
    }
  }
  sim.flags = flags;
}

// Continuous collision of the fast bodies finalize listed, one fast body per lane,
// then island sleep for worlds with sleep enabled. Worlds with neither return at
// once. part 0: the whole pass, one workgroup per world (blockIdx.x). part 1: the
// non-bullet sweeps only, workgroup blockIdx.x of gridDim.x taking every
// gridDim.x-th stride of the fast list of world blockIdx.y (non-bullet fast bodies
// sweep against static shapes only, so they are independent). part 2: the bullets
// and the sleep phase after a part-1 launch, one workgroup per world.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void continuousKernel(BatchView b, const int* worldList, int part) {
  const int wi = part == 1 ? (worldList ? worldList[blockIdx.y] : blockIdx.y)
                           : (worldList ? worldList[blockIdx.x] : blockIdx.x);
  WorldHeader* h = b.worlds + wi;
  const int nFast = h->fastCount;
  const bool sleep = h->sleepPending != 0;
  if (nFast == 0 && (!sleep || part == 1)) return;
  if (part == 1 && (int)blockIdx.x * BLOCK >= nFast) return;
  WorldPtrs w = worldPtrs(b, h);
  const int ns = h->shapeCount;
  const int* fastList = w.scratch;  // written by finalize, nFast <= shapeCount
  const int* staticList = b.staticShapes + h->shapeOffset;
  const int staticCount = h->staticCount;
  (void)ns;
  // every fast body tests every static proxy: their boxes once per world in LDS
  __shared__ StaticProxies statics;
  for (int k = threadIdx.x; k < min(staticCount, kMaxStagedStatics); k += BLOCK) {
    const Shape& t = w.shapes[staticList[k]];
    statics.fat[k] = t.fat;
    statics.visible[k] = ((t.catLo | t.catHi) != 0 && t.proxyKey >= 0) ? 1 : 0;
  }
  __syncthreads();
  if (part != 2) {
    const int first = part == 1 ? (int)blockIdx.x * BLOCK : 0;
    const int stride = part == 1 ? (int)gridDim.x * BLOCK : BLOCK;
    for (int k = first + threadIdx.x; k < nFast; k += stride) {
      int i = fastList[k];
      if ((w.sims[i].flags & BF_BULLET) == 0) solveContinuous(w, i, staticList, staticCount, false, statics);
    }
    if (part == 1) return;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nFast; k += BLOCK) {
    int i = fastList[k];
    if (w.sims[i].flags & BF_BULLET) solveContinuous(w, i, staticList, staticCount, true, statics);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    h->fastCount = 0;
    if (sleep) sleepPhase(w);
  }
}

// Island commands of the body and world API, one world, thread 0:
//   1  b2WakeBody: wake the sleeping set of body `arg` (S/body.c:332-344)
//   2  b2World_EnableSleeping(false): wake every sleeping set in set order (S/world.c:1710-1737)
//   3  b2Body_SetAwake(false): split the body's island if it lost constraints, then
//      try to put it to sleep (S/body.c:1483-1515)
__global__ void islandCommandKernel(BatchView b, int world, int op, int arg) {
  if (threadIdx.x != 0) return;
  WorldHeader* h = b.worlds + world;
  WorldPtrs w = worldPtrs(b, h);
  IslandCtx c = w.isl;
  if (op == 1) {
    if (arg >= 0 && arg < h->bodyCount && w.sims[arg].setIndex >= kFirstSleepingSet) wakeIslandSet(w, w.sims[arg].islandId);
  } else if (op == 2) {
    for (int set = kFirstSleepingSet; set < h->setCount; ++set) {
      for (int i = 0; i < h->islandCount; ++i) {
        if (c.islands[i].id == i && c.islands[i].setIndex == set) {
          wakeIslandSet(w, i);
          break;
        }
      }
    }
  } else if (op == 3) {
    if (arg >= 0 && arg < h->bodyCount && w.sims[arg].setIndex == kAwakeSet && w.sims[arg].islandId >= 0) {
      if (c.islands[w.sims[arg].islandId].removeCount > 0) splitIsland(c, w.sims[arg].islandId);
      trySleepIsland(w, w.sims[arg].islandId);
    }
  }
}

// ---------------------------------------------------------------------------- queries

constexpr int kQueryLanes = 64;  // one query per lane
constexpr int kQueryStack = 64;  // per-lane traversal stack (entries): a tree of height h needs h + 1

__device__ __forceinline__ Vec2 boxCenter(const AABB& a) {  // b2AABB_Center
  return {0.5f * (a.lo.x + a.hi.x), 0.5f * (a.lo.y + a.hi.y)};
}
__device__ __forceinline__ Vec2 boxExtents(const AABB& a) {  // b2AABB_Extents
  return {0.5f * (a.hi.x - a.lo.x), 0.5f * (a.hi.y - a.lo.y)};
}

// The leaves the last step enlarged (SF_TREE_PENDING), applied to the world's
// kinematic and dynamic trees as b2DynamicTree_EnlargeProxy does
// (S/dynamic_tree.c:827-866): the leaf takes the fat box, its ancestors grow to hold
// it and are marked enlarged. The step's tree launch applies them at the start of
// the next step; queries between steps walk the trees and need them now. The
// boxes and marks do not depend on the order of the leaves, and the next tree
// launch rebuilds the marked nodes as it would have. Thread 0 of one workgroup per
// world walks up from each pending leaf.
__global__ __launch_bounds__(64) void enlargePendingKernel(BatchView b) {
  const int world = blockIdx.x;
  WorldHeader* h = b.worlds + world;
  Shape* shapes = b.shapes + h->shapeOffset;
  const int ns = h->shapeCount;
  bool any = false;
  for (int k = threadIdx.x; k < ns; k += 64) any |= (shapes[k].flags & SF_TREE_PENDING) != 0;
  if (!__syncthreads_or(any) || threadIdx.x != 0) return;
  for (int k = 0; k < ns; ++k) {
    Shape& sh = shapes[k];
    if (!(sh.flags & SF_TREE_PENDING)) continue;
    const int t = sh.proxyKey & 3;
    if (sh.proxyKey < 0 || t == kStatic) continue;
    TreeHeader* hd = b.treeHdrs + 3L * world + t;
    TreeNode* n = b.treeNodes + hd->offset;
    const int F = hd->fresh;
    const int leaf = sh.proxyKey >> 2;
    if (leaf < 0 || leaf >= F) continue;  // left to the tree launch, which flags the tree
    sh.flags &= ~SF_TREE_PENDING;
    const AABB box = sh.fat;
    n[leaf].box = box;
    int p = n[leaf].parent;
    for (int g = 0; p >= 0 && p < F && g < F; ++g) {
      const bool changed = aabbEnlarge(&n[p].box, box);
      n[p].flags |= TN_ENLARGED;
      p = n[p].parent;
      if (!changed) break;
    }
    for (int g = 0; p >= 0 && p < F && g < F; ++g) {
      if (n[p].flags & TN_ENLARGED) break;
      n[p].flags |= TN_ENLARGED;
      p = n[p].parent;
    }
  }
}

// Per-lane traversal stack in LDS, entries interleaved across the wavefront's lanes.
struct LaneStack {
  int* base;
  int count;
  bool overflow;
  __device__ __forceinline__ void push(int v) {
    if (count < kQueryStack) {
      base[count * kQueryLanes] = v;
      ++count;
    } else {
      overflow = true;
    }
  }
  __device__ __forceinline__ int pop() { return base[--count * kQueryLanes]; }
};

// One query of the batch query ABI per lane, walking the world's trees the way
// the reference's world queries do, so hits come out in the reference's order:
//   QK_RAY_CLOSEST  b2World_CastRayClosest (S/world.c:2277-2308): each tree in turn
//                   (static, kinematic, dynamic) by b2DynamicTree_RayCast
//                   (S/dynamic_tree.c:1172-1290: the nearer child first, the segment
//                   shortened by every accepted hit) with b2RayCastClosestFcn (initial
//                   overlaps ignored; the last hit at the smallest fraction is kept)
//   QK_RAY_ALL      b2World_CastRay (S/world.c:2222-2254) with a callback that returns
//                   1: every hit, initial overlaps included, in traversal order
//   QK_OVERLAP_AABB b2World_OverlapAABB (S/world.c:2071-2105): each tree's
//                   b2DynamicTree_Query (S/dynamic_tree.c:1114-1170, child2 popped
//                   first), every shape whose fat AABB overlaps the box
// counts[q] receives the number of results (which may exceed cap: only the first cap
// are written to hits[q * cap ...]); -1 for a world index outside the batch; -2 if a
// tree was deeper than the traversal stack (never for trees of realistic height).
__device__ void runQuery(const BatchView& b, const Query& q, QueryHit* out, int* count, int cap, LaneStack& st) {
  if (q.world < 0 || q.world >= b.worldCount) {
    *count = -1;
    return;
  }
  const WorldHeader* h = b.worlds + q.world;
  const Shape* shapes = b.shapes + h->shapeOffset;
  const ShapeGeom* geoms = b.geoms + h->shapeOffset;
  const BodySim* sims = b.sims + h->bodyOffset;
  const TreeHeader* th = b.treeHdrs + 3L * q.world;
  int total = 0;
  if (q.kind == QK_OVERLAP_AABB) {
    const AABB box = {{q.p[0], q.p[1]}, {q.p[2], q.p[3]}};
    for (int t = kStatic; t <= kDynamic; ++t) {
      const TreeHeader& hd = th[t];
      if (hd.nodeCount == 0 || hd.root < 0) continue;
      const TreeNode* n = b.treeNodes + hd.offset;
      st.count = 0;
      st.push(hd.root);
      while (st.count > 0) {
        const TreeNode& node = n[st.pop()];
        if (!aabbOverlaps(node.box, box)) continue;
        if (node.flags & TN_LEAF) {
          const int s = node.child1;
          if (!shouldQuery(shapes[s], q.catLo, q.catHi, q.maskLo, q.maskHi)) continue;
          if (total < cap) {
            QueryHit r;
            r.shape = s;
            r.px = r.py = r.nx = r.ny = r.fraction = 0.0f;
            out[total] = r;
          }
          ++total;
        } else {
          st.push(node.child1);
          st.push(node.child2);
        }
      }
    }
    *count = st.overflow ? -2 : total;
    return;
  }
  if (q.kind != QK_RAY_CLOSEST && q.kind != QK_RAY_ALL) {
    *count = 0;
    return;
  }
  const bool closest = q.kind == QK_RAY_CLOSEST;
  const Vec2 p1 = {q.p[0], q.p[1]}, d = {q.p[2], q.p[3]};
  const Vec2 r = normalize(d);
  const Vec2 v = crossSV(1.0f, r);
  const Vec2 absV = {absf(v.x), absf(v.y)};
  float inputMax = 1.0f;     // the world query's input.maxFraction, carried from tree to tree
  float worldFraction = 1.0f;
  QueryHit best;
  bool haveBest = false;
  for (int t = kStatic; t <= kDynamic; ++t) {
    const TreeHeader& hd = th[t];
    if (hd.nodeCount != 0 && hd.root >= 0) {
      const TreeNode* n = b.treeNodes + hd.offset;
      float maxFraction = inputMax;
      Vec2 p2 = mulAdd(p1, maxFraction, d);
      AABB seg = {vmin(p1, p2), vmax(p1, p2)};
      st.count = 0;
      st.push(hd.root);
      bool terminated = false;
      while (st.count > 0 && !terminated) {
        const TreeNode& node = n[st.pop()];
        if (!aabbOverlaps(node.box, seg)) continue;
        const Vec2 c = boxCenter(node.box), e = boxExtents(node.box);
        const float term1 = absf(dot(v, sub(p1, c)));
        const float term2 = dot(absV, e);
        if (term2 < term1) continue;
        if (node.flags & TN_LEAF) {
          // RayCastCallback, S/world.c:2191-2220
          const int s = node.child1;
          const Shape& sh = shapes[s];
          float value = maxFraction;
          if (shouldQuery(sh, q.catLo, q.catHi, q.maskLo, q.maskHi)) {
            const RayInput in = {p1, d, maxFraction};
            const CastOutput o = rayCastShape(in, geoms[s], sims[sh.body].xf);
            if (o.hit) {
              float fraction;
              QueryHit rec;
              rec.shape = s;
              rec.px = o.point.x;
              rec.py = o.point.y;
              rec.nx = o.normal.x;
              rec.ny = o.normal.y;
              rec.fraction = o.fraction;
              if (closest) {
                // b2RayCastClosestFcn, S/world.c:2258-2275
                if (o.fraction == 0.0f) {
                  fraction = -1.0f;
                } else {
                  best = rec;
                  haveBest = true;
                  fraction = o.fraction;
                }
              } else {
                if (total < cap) out[total] = rec;
                ++total;
                fraction = 1.0f;
              }
              if (0.0f <= fraction && fraction <= 1.0f) worldFraction = fraction;
              value = fraction;
            }
          }
          if (value == 0.0f) {
            terminated = true;
          } else if (0.0f < value && value <= maxFraction) {
            maxFraction = value;
            p2 = mulAdd(p1, maxFraction, d);
            seg = {vmin(p1, p2), vmax(p1, p2)};
          }
        } else {
          const Vec2 c1 = boxCenter(n[node.child1].box), c2 = boxCenter(n[node.child2].box);
          if (distanceSq(c1, p1) < distanceSq(c2, p1)) {
            st.push(node.child2);
            st.push(node.child1);
          } else {
            st.push(node.child1);
            st.push(node.child2);
          }
        }
      }
    }
    if (worldFraction == 0.0f) break;
    inputMax = worldFraction;
  }
  if (closest) {
    if (haveBest && cap > 0) out[0] = best;
    total = haveBest ? 1 : 0;
  }
  *count = st.overflow ? -2 : total;
}

__global__ __launch_bounds__(kQueryLanes) void queryKernel(BatchView b, const Query* qs, const int* order, int n,
                                                           QueryHit* hits, int* counts, int cap) {
  __shared__ int stackMem[kQueryStack * kQueryLanes];
  const int k = blockIdx.x * kQueryLanes + threadIdx.x;
  if (k >= n) return;
  const int qi = order ? order[k] : k;
  LaneStack st = {stackMem + threadIdx.x, 0, false};
  runQuery(b, qs[qi], hits + (long)qi * cap, counts + qi, cap, st);
}

// Queries grouped by world (a counting sort into `order`), so that the lanes of a
// wavefront walk the same trees: the count of each world's queries, the worlds'
// first slots (one workgroup scan), then each query's slot.
__global__ void queryCountKernel(const Query* qs, int n, int* worldCount, int W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int w = qs[i].world;
  atomicAdd(worldCount + ((w >= 0 && w < W) ? w : W), 1);
}

__global__ __launch_bounds__(1024) void queryScanKernel(int* worldCount, int W) {
  __shared__ int sums[1024 / 64];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base <= W; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i <= W ? worldCount[i] : 0;
    int total;
    const int incl = blockInclusiveAdd<1024>(v, sums, &total);
    if (i <= W) worldCount[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += total;
    __syncthreads();
  }
}

__global__ void queryScatterKernel(const Query* qs, int n, int* worldStart, int W, int* order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int w = qs[i].world;
  order[atomicAdd(worldStart + ((w >= 0 && w < W) ? w : W), 1)] = i;
}

// b2World_Explode, S/world.c:2722-2747: the dynamic tree's b2DynamicTree_Query over
// the explosion's box in depth-first order, ExplosionCallback (S/world.c:2640-2720)
// on each leaf in turn: the distance from the shape to the explosion, the body woken
// (its sleeping set, S/body.c:332-344), and the impulse added to an awake body's
// velocities. Thread 0 walks; one workgroup per explosion.
__global__ __launch_bounds__(kQueryLanes) void explodeKernel(BatchView b, const Query* qs) {
  __shared__ int stackMem[kQueryStack * kQueryLanes];
  const Query q = qs[blockIdx.x];
  if (threadIdx.x != 0 || q.kind != QK_EXPLODE || q.world < 0 || q.world >= b.worldCount) return;
  WorldHeader* h = b.worlds + q.world;
  WorldPtrs w = worldPtrs(b, h);
  const TreeHeader& hd = b.treeHdrs[3L * q.world + kDynamic];
  if (hd.nodeCount == 0 || hd.root < 0) return;
  const TreeNode* n = b.treeNodes + hd.offset;
  const Vec2 position = {q.p[0], q.p[1]};
  const float radius = q.p[2], falloff = q.p[3], impulsePerLength = q.p[4];
  AABB box;
  box.lo.x = position.x - (radius + falloff);
  box.lo.y = position.y - (radius + falloff);
  box.hi.x = position.x + (radius + falloff);
  box.hi.y = position.y + (radius + falloff);
  const Xf identity = {{0.0f, 0.0f}, {1.0f, 0.0f}};
  LaneStack st = {stackMem, 0, false};
  st.push(hd.root);
  while (st.count > 0) {
    const TreeNode& node = n[st.pop()];
    if (!aabbOverlaps(node.box, box)) continue;
    if (!(node.flags & TN_LEAF)) {
      st.push(node.child1);
      st.push(node.child2);
      continue;
    }
    const int s = node.child1;
    const Shape& sh = w.shapes[s];
    // the tree's category test against the explosion's mask
    if (((sh.catLo & q.maskLo) | (sh.catHi & q.maskHi)) == 0u) continue;
    const int body = sh.body;
    const Xf xf = w.sims[body].xf;
    const ShapeGeom& g = w.geoms[s];
    Proxy pa = makeShapeProxy(g);
    Proxy pb;
    pb.count = 1;
    pb.p[0] = position;
    pb.radius = 0.0f;
    SimplexCache cache;
    cache.count = 0;
    const DistanceOut dist = shapeDistanceOut(cache, pa, pb, xf, identity, true);
    if (dist.distance > radius + falloff) continue;
    if (w.sims[body].setIndex >= kFirstSleepingSet && w.sims[body].islandId >= 0) wakeIslandSet(w, w.sims[body].islandId);
    if (w.sims[body].setIndex != kAwakeSet) continue;
    Vec2 closestPoint = dist.pointA;
    if (dist.distance == 0.0f) closestPoint = transformPoint(xf, shapeCentroid(g));
    Vec2 direction = sub(closestPoint, position);
    if (lengthSq(direction) > 100.0f * kFltEpsilon * kFltEpsilon) {
      direction = normalize(direction);
    } else {
      direction = {1.0f, 0.0f};
    }
    const Vec2 localLine = invRotate(xf.q, leftPerp(direction));
    const float perimeter = projectedPerimeter(g, localLine);
    float scale = 1.0f;
    if (dist.distance > radius && falloff > 0.0f) scale = clampf((radius + falloff - dist.distance) / falloff, 0.0f, 1.0f);
    const float magnitude = impulsePerLength * perimeter * scale;
    const Vec2 impulse = mulSV(magnitude, direction);
    const BodySim& sim = w.sims[body];
    BodyState& bs = w.states[body];
    bs.v = mulAdd(bs.v, sim.invMass, impulse);
    bs.w += sim.invI * cross(sub(closestPoint, sim.center), impulse);
  }
  if (st.overflow) h->error |= ERR_TREE;
}

// ---------------------------------------------------------------------------- broad-phase trees

// The tree work of b2Collide for one world per workgroup (S/world.c:447-459, 494-501):
// for the dynamic and the kinematic tree, the leaves the last step enlarged
// (S/solver.c:1848-1903, 1959-1985), the partial rebuild (S/broad_phase.c:488-492) and
// every proxy's rank in the tree's query order for the next pair update
// (rebuildTree). The work arrays live in LDS when the tree fits, else in the
// tree's HBM work region.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) treeKernel(BatchView b, const int* worldList, int ldsBytes) {
  extern __shared__ __align__(16) int treeLds[];
  const int world = worldList[blockIdx.x];
  const WorldHeader* h = b.worlds + world;
  Shape* shapes = b.shapes + h->shapeOffset;
  int* ranks = b.treeRanks + h->shapeOffset;
  TreeHeader* th = b.treeHdrs + 3L * world;
  for (int t = kDynamic; t >= kKinematic; --t) {
    TreeHeader* hd = th + t;
    const int F = hd->fresh, P = hd->proxyCount;
    if (P == 0) continue;
    // two inlined copies: the work arrays' address space stays known (LDS or HBM)
    if (treeWorkInts(F, P) * 4 <= ldsBytes) {
      rebuildTree<BLOCK>(hd, b.treeNodes + hd->offset, shapes, h->shapeCount, t, ranks, treeWorkAt(treeLds, F, P));
    } else {
      rebuildTree<BLOCK>(hd, b.treeNodes + hd->offset, shapes, h->shapeCount, t, ranks,
                         treeWorkAt(b.treeWork + treeWorkOffset(hd->offset, world, t), F, P));
    }
    __syncthreads();
  }
}

static void launchTree(const BatchView* b, const int* worldList, int nWorlds, hipStream_t stream,
                       const LaunchConfig* cfg) {
  if (cfg->treeBlock >= 1024) {
    hipLaunchKernelGGL(treeKernel<1024>, dim3(nWorlds), dim3(1024), cfg->treeLds, stream, *b, worldList, cfg->treeLds);
  } else if (cfg->treeBlock >= 256) {
    hipLaunchKernelGGL(treeKernel<256>, dim3(nWorlds), dim3(256), cfg->treeLds, stream, *b, worldList, cfg->treeLds);
  } else if (cfg->treeBlock >= 128) {
    hipLaunchKernelGGL(treeKernel<128>, dim3(nWorlds), dim3(128), cfg->treeLds, stream, *b, worldList, cfg->treeLds);
  } else {
    hipLaunchKernelGGL(treeKernel<64>, dim3(nWorlds), dim3(64), cfg->treeLds, stream, *b, worldList, cfg->treeLds);
  }
}

template <int BLOCK>
static int launchBlock(const BatchView* b, const StepParams* sp, const int* worldList, int nWorlds, hipStream_t stream,
                       const LaunchConfig* cfg) {
  if ((sp->phases & PH_PAIRS) && cfg->sweepChunks > 0 && nWorlds <= 65535) {
    // large worlds: sort, sweep over many workgroups, then create, collide, state
    StepParams a = *sp;
    a.phases = PH_PAIRS | PH_PAIRS_PREP | PH_NO_COUNT;
    hipLaunchKernelGGL(collideKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->collideLds, stream, *b, a, worldList,
                       cfg->collideLds);
    hipLaunchKernelGGL(sweepKernel, dim3(cfg->sweepChunks, nWorlds), dim3(256), 0, stream, *b, worldList);
    StepParams c = *sp;
    c.phases = (sp->phases & ~(PH_PAIRS | PH_COLLIDE | PH_STATE)) | PH_PAIRS_FINISH | PH_NO_COUNT;
    hipLaunchKernelGGL(collideKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->collideLds, stream, *b, c, worldList,
                       cfg->collideLds);
    if (sp->phases & PH_COLLIDE) {
      hipLaunchKernelGGL(narrowKernel, dim3(cfg->sweepChunks * 2, nWorlds), dim3(256), 0, stream, *b, worldList);
    }
    StepParams d = *sp;
    d.phases = (sp->phases & ~PH_PAIRS) | PH_CONTACTS_DONE;
    hipLaunchKernelGGL(collideKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->collideLds, stream, *b, d, worldList,
                       cfg->collideLds);
  } else if (sp->phases & (PH_PAIRS | PH_COLLIDE | PH_STATE)) {
    hipLaunchKernelGGL(collideKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->collideLds, stream, *b, *sp, worldList,
                       cfg->collideLds);
  }
  if (((sp->phases & PH_COLLIDE) && !(sp->phases & PH_NO_TREE)) || (sp->phases & PH_TREE)) {
    launchTree(b, worldList, nWorlds, stream, cfg);
  }
  if (cfg->wideChunks > 0 && (sp->phases & PH_SOLVE) && nWorlds <= 65535) {
    // large worlds: islands and colour buckets in one workgroup per world, then
    // every stage over many workgroups, then finalize in one workgroup per world
    StepParams p = *sp;
    p.phases = PH_SOLVE | PH_SOLVE_PREP | ((sp->phases & PH_FINALIZE) ? PH_NO_COUNT : 0);
    hipLaunchKernelGGL(solveKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), 0, stream, *b, p, worldList, 0);
    const dim3 grid(cfg->wideChunks, nWorlds);
    auto stage = [&](int st, int color) {
      hipLaunchKernelGGL(wideStageKernel, grid, dim3(256), 0, stream, *b, *sp, worldList, st, color);
    };
    auto colors = [&](int st) {
      stage(st, kOverflowIndex);
      for (int c = 0; c < kOverflowIndex; ++c) stage(st, c);
    };
    stage(WS_PREPARE, 0);
    for (int sub = 0; sub < sp->subSteps; ++sub) {
      stage(WS_INTEGRATE_VEL, 0);
      colors(WS_WARM);
      colors(WS_SOLVE);
      stage(WS_INTEGRATE_POS, 0);
      colors(WS_RELAX);
    }
    colors(WS_RESTITUTION);
    stage(WS_STORE, 0);
    if (sp->phases & PH_FINALIZE) {
      StepParams f = *sp;
      f.phases = sp->phases & ~PH_SOLVE;
      hipLaunchKernelGGL(solveKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->solveLds, stream, *b, f, worldList,
                         cfg->solveLds);
    }
  } else if (sp->phases & (PH_SOLVE | PH_FINALIZE)) {
    hipLaunchKernelGGL(solveKernel<BLOCK>, dim3(nWorlds), dim3(BLOCK), cfg->solveLds, stream, *b, *sp, worldList,
                       cfg->solveLds);
  }
  bool ccd = (sp->phases & PH_CONTINUOUS) || ((sp->phases & PH_FINALIZE) && !(sp->phases & PH_NO_CONTINUOUS));
  if (ccd && cfg->continuous) {
    if (cfg->continuousChunks > 0 && nWorlds <= 65535) {
      hipLaunchKernelGGL(continuousKernel<kContinuousLanes>, dim3(cfg->continuousChunks, nWorlds),
                         dim3(kContinuousLanes), 0, stream, *b, worldList, 1);
      hipLaunchKernelGGL(continuousKernel<kContinuousLanes>, dim3(nWorlds), dim3(kContinuousLanes), 0, stream, *b,
                         worldList, 2);
    } else {
      hipLaunchKernelGGL(continuousKernel<kContinuousLanes>, dim3(nWorlds), dim3(kContinuousLanes), 0, stream, *b,
                         worldList, 0);
    }
  }
  return (int)hipGetLastError();
}

}  // namespace f2d

// Zero n ints (a fill kernel, not hipMemsetAsync, so that callers may capture the
// launches in a graph).
__global__ void zeroIntsKernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// The pending leaf enlargements of every world applied to its trees (before queries
// walk them).
extern "C" int f2d_launch_enlarge_pending(const f2d::BatchView* b, hipStream_t stream) {
  if (b->worldCount <= 0) return 0;
  hipLaunchKernelGGL(f2d::enlargePendingKernel, dim3((unsigned)b->worldCount), dim3(64), 0, stream, *b);
  return (int)hipGetLastError();
}

// Queries, one per lane. group (optional): n + worldCount + 1 ints of scratch for
// grouping the queries by world first; null runs them in the given order. The
// results do not depend on the grouping.
extern "C" int f2d_launch_queries(const f2d::BatchView* b, int* group, const f2d::Query* q, int n,
                                  f2d::QueryHit* hits, int* counts, int cap, hipStream_t stream) {
  if (n <= 0) return 0;
  if (f2d_launch_enlarge_pending(b, stream) != 0) return -1;
  const int W = b->worldCount;
  int* order = nullptr;
  if (group) {
    order = group;
    int* worldCount = group + n;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(zeroIntsKernel, dim3((unsigned)((W + 1 + 255) / 256)), dim3(256), 0, stream, worldCount, W + 1);
    hipLaunchKernelGGL(f2d::queryCountKernel, dim3(nb), dim3(256), 0, stream, q, n, worldCount, W);
    hipLaunchKernelGGL(f2d::queryScanKernel, dim3(1), dim3(1024), 0, stream, worldCount, W);
    hipLaunchKernelGGL(f2d::queryScatterKernel, dim3(nb), dim3(256), 0, stream, q, n, worldCount, W, order);
  }
  const unsigned groups = (unsigned)((n + f2d::kQueryLanes - 1) / f2d::kQueryLanes);
  hipLaunchKernelGGL(f2d::queryKernel, dim3(groups), dim3(f2d::kQueryLanes), 0, stream, *b, q, order, n, hits, counts,
                     cap);
  return (int)hipGetLastError();
}

extern "C" int f2d_launch_explosions(const f2d::BatchView* b, const f2d::Query* q, int n, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(f2d::explodeKernel, dim3(n), dim3(f2d::kQueryLanes), 0, stream, *b, q);
  return (int)hipGetLastError();
}

extern "C" int f2d_launch_island_command(const f2d::BatchView* b, int world, int op, int arg, hipStream_t stream) {
  hipLaunchKernelGGL(f2d::islandCommandKernel, dim3(1), dim3(64), 0, stream, *b, world, op, arg);
  return (int)hipGetLastError();
}

extern "C" int f2d_launch_step(const f2d::BatchView* b, const f2d::StepParams* sp, const int* worldList, int nWorlds,
                               hipStream_t stream, const f2d::LaunchConfig* cfg) {
  if (nWorlds <= 0) return 0;
  switch (cfg->block) {
    case 64:
      return f2d::launchBlock<64>(b, sp, worldList, nWorlds, stream, cfg);
    case 128:
      return f2d::launchBlock<128>(b, sp, worldList, nWorlds, stream, cfg);
    case 512:
      return f2d::launchBlock<512>(b, sp, worldList, nWorlds, stream, cfg);
    case 1024:
      return f2d::launchBlock<1024>(b, sp, worldList, nWorlds, stream, cfg);
    default:
      return f2d::launchBlock<256>(b, sp, worldList, nWorlds, stream, cfg);
  }
}

// ---------------------------------------------------------------------------- allocation-free queries

#include <mutex>

#include "world_store.hpp"

namespace f2d {

// The query grouping scratch of the allocation-free query entry point (the query
// order, then a count and first slot per world): file-scope device memory, one copy
// per GPU, held until the process ends. Calls with more queries plus worlds than
// it holds are refused.
constexpr long kStaticGroupCapacity = 1L << 20;  // ints (4 MiB)
static __device__ __align__(256) int scratch[kStaticGroupCapacity];

constexpr int kMaxDevices = 64;
static std::mutex g_scratchLock;
static int* g_scratchAt[kMaxDevices];  // the array's device address, per device

// The array's address on the current device, looked up once per device. In host
// code `scratch` names a host-side stand-in, never the device memory.
static int* staticScratch(std::string& err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    err = "no current HIP device";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratchLock);
  if (!g_scratchAt[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratchAt[dev] = static_cast<int*>(p);
  }
  return g_scratchAt[dev];
}

}  // namespace f2d

// f2dg_query_device with the query grouping in the static array instead of a buffer
// the library allocates: no allocation of any kind on this path. The batch must be
// uploaded already (f2dg_upload or a step) with no host edits pending, and hits and
// counts must be device memory. Calls on one GPU share the array: calls that
// overlap on one GPU race.
extern "C" int f2dg_query_device_static(void* h, const void* queries, int n, void* hits, void* counts, int cap) {
  f2d::Batch* b = static_cast<f2d::Batch*>(h);
  if (!b || n < 0 || cap < 1) return -1;
  if (!b->committed || b->hostDirty) {
    b->lastError = "f2dg_query_device_static: upload the batch first (it allocates nothing itself)";
    return -1;
  }
  const long need = (long)n + (long)b->worlds.size() + 1;
  if (need > f2d::kStaticGroupCapacity) {
    b->lastError = "f2dg_query_device_static: " + std::to_string(n) + " queries over " +
                   std::to_string(b->worlds.size()) + " worlds need " + std::to_string(need) +
                   " ints of grouping scratch, more than the static array holds (" +
                   std::to_string(f2d::kStaticGroupCapacity) + ")";
    return -1;
  }
  int* group = f2d::staticScratch(b->lastError);
  if (!group) return -1;
  if (n == 0) return 0;
  b->deviceNewer = true;  // pending tree enlargements applied
  return f2d_launch_queries(&b->dev.view, group, static_cast<const f2d::Query*>(queries), n,
                            static_cast<f2d::QueryHit*>(hits), static_cast<int*>(counts), cap, b->stream);
}
