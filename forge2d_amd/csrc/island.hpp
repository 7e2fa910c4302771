// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Island graph bookkeeping: union-find linking of touching contacts and joints,
// merging of awake islands, and the depth-first split of an island that lost
// constraints (S/island.c). The same code runs on the host (joint creation) and
// inside the step kernels (one thread of the world's workgroup): it is serial
// integer work over a world's linked lists.
//
// Island records, the id pools and the awake set's orders live in the world's
// island region (IslandRegion); contacts keep their island links in a parallel
// int4 array (island id, prev, next, solver set), bodies and joints in their own
// records.
#pragma once

#include "layout.hpp"

namespace f2d {

struct IslandCtx {
  WorldHeader* h;
  Island* islands;
  int* islandFree;
  int* awakeIslands;
  int* setFree;
  int* awakeBodies;
  int* moveOrder;
  uint32_t* awakeBits;
  int* stack;
  int* seeds;
  BodySim* sims;
  Contact* contacts;
  int4* links;
  Joint* joints;
};

F2D_HD IslandCtx islandCtx(WorldHeader* h, int* region, BodySim* sims, Contact* contacts, int4* links,
                           Joint* joints) {
  IslandRegion r = islandRegion(h->islandCap, h->bodyCount);
  IslandCtx c;
  c.h = h;
  c.islands = reinterpret_cast<Island*>(region + r.islands);
  c.islandFree = region + r.islandFree;
  c.awakeIslands = region + r.awakeIslands;
  c.setFree = region + r.setFree;
  c.awakeBodies = region + r.awakeBodies;
  c.moveOrder = region + r.moveOrder;
  c.awakeBits = reinterpret_cast<uint32_t*>(region + r.awakeBits);
  c.stack = region + r.stack;
  c.seeds = region + r.seeds;
  c.sims = sims;
  c.contacts = contacts;
  c.links = links;
  c.joints = joints;
  return c;
}

// b2AllocId, S/id_pool.c:19-37: pop the free list, else the next fresh id
F2D_HD int poolAlloc(int* freeArr, int& freeCount, int& next) {
  if (freeCount > 0) return freeArr[--freeCount];
  return next++;
}

// b2CreateIsland, S/island.c:18-55, for the awake set. Returns -1 when the region is full.
F2D_HD int createIsland(IslandCtx& c) {
  WorldHeader* h = c.h;
  if (h->islandFreeCount == 0 && h->islandCount >= h->islandCap) return -1;
  int id = poolAlloc(c.islandFree, h->islandFreeCount, h->islandCount);
  Island& is = c.islands[id];
  is.id = id;
  is.setIndex = kAwakeSet;
  is.localIndex = h->awakeIslandCount;
  is.headBody = is.tailBody = -1;
  is.bodyCount = 0;
  is.headContact = is.tailContact = -1;
  is.contactCount = 0;
  is.headJoint = is.tailJoint = -1;
  is.jointCount = 0;
  is.parent = -1;
  is.removeCount = 0;
  c.awakeIslands[h->awakeIslandCount++] = id;
  return id;
}

// b2DestroyIsland, S/island.c:57-83: remove-swap from the awake set's island
// order; the record keeps its other fields
F2D_HD void destroyIsland(IslandCtx& c, int id) {
  WorldHeader* h = c.h;
  if (h->splitIslandId == id) h->splitIslandId = -1;
  Island& is = c.islands[id];
  int li = is.localIndex, last = --h->awakeIslandCount;
  if (li != last) {
    int moved = c.awakeIslands[last];
    c.awakeIslands[li] = moved;
    c.islands[moved].localIndex = li;
  }
  is.id = -1;
  is.setIndex = -1;
  is.localIndex = -1;
  c.islandFree[h->islandFreeCount++] = id;
}

// b2AddContactToIsland, S/island.c:85-112 (prepend)
F2D_HD void addContactToIsland(IslandCtx& c, int islandId, int cid) {
  Island& is = c.islands[islandId];
  int4& l = c.links[cid];
  if (is.headContact != -1) {
    l.z = is.headContact;
    c.links[is.headContact].y = cid;
  }
  is.headContact = cid;
  if (is.tailContact == -1) is.tailContact = is.headContact;
  is.contactCount += 1;
  l.x = islandId;
}

// b2AddJointToIsland, S/island.c:264-287 (prepend)
F2D_HD void addJointToIsland(IslandCtx& c, int islandId, int jid) {
  Island& is = c.islands[islandId];
  Joint& j = c.joints[jid];
  if (is.headJoint != -1) {
    j.islandNext = is.headJoint;
    c.joints[is.headJoint].islandPrev = jid;
  }
  is.headJoint = jid;
  if (is.tailJoint == -1) is.tailJoint = is.headJoint;
  is.jointCount += 1;
  j.islandId = islandId;
}

// Union-find root with path halving: every island left on the way up points to
// its grandparent (the walk of S/island.c:150-196 and :547-566)
F2D_HD int islandRoot(IslandCtx& c, int id) {
  Island* is = &c.islands[id];
  // parents form a forest; the walk is bounded by the island count all the same
  for (int guard = c.h->islandCount; is->parent != -1 && guard > 0; --guard) {
    int pid = is->parent;
    Island* parent = &c.islands[pid];
    if (parent->parent != -1) is->parent = parent->parent;
    is = parent;
    id = pid;
  }
  return id;
}

// The island part of b2LinkContact, S/island.c:148-218 (the caller wakes sleeping
// sets first); iA, iB: the islands of the contact's bodies
F2D_HD void linkContactIslandIds(IslandCtx& c, int cid, int iA, int iB) {
  if (iA == iB) {
    addContactToIsland(c, iA, cid);
    return;
  }
  int rA = iA != -1 ? islandRoot(c, iA) : -1;
  int rB = iB != -1 ? islandRoot(c, iB) : -1;
  if (rA != -1 && rB != -1 && rA != rB) {
    c.islands[rB].parent = rA;
    c.h->mergePending = 1;
  }
  addContactToIsland(c, rA != -1 ? rA : rB, cid);
}

F2D_HD_COLD void linkContactIslands(IslandCtx& c, int cid) {
  const Contact& ct = c.contacts[cid];
  linkContactIslandIds(c, cid, c.sims[ct.bodyA].islandId, c.sims[ct.bodyB].islandId);
}

// b2UnlinkContact, S/island.c:221-262
F2D_HD_COLD void unlinkContact(IslandCtx& c, int cid) {
  int4& l = c.links[cid];
  Island& is = c.islands[l.x];
  if (l.y != -1) c.links[l.y].z = l.z;
  if (l.z != -1) c.links[l.z].y = l.y;
  if (is.headContact == cid) is.headContact = l.z;
  if (is.tailContact == cid) is.tailContact = l.y;
  is.contactCount -= 1;
  is.removeCount += 1;
  l.x = l.y = l.z = -1;
}

// The island part of b2LinkJoint, S/island.c:289-385, for joints created between
// awake or static bodies (creation time; the caller merges islands afterwards)
F2D_HD_COLD void linkJointIslands(IslandCtx& c, int jid) {
  const Joint& j = c.joints[jid];
  int iA = c.sims[j.bodyA].islandId, iB = c.sims[j.bodyB].islandId;
  if (iA == iB) {
    addJointToIsland(c, iA, jid);
    return;
  }
  int rA = iA != -1 ? islandRoot(c, iA) : -1;
  int rB = iB != -1 ? islandRoot(c, iB) : -1;
  if (rA != -1 && rB != -1 && rA != rB) {
    c.islands[rB].parent = rA;
    c.h->mergePending = 1;
  }
  addJointToIsland(c, rA != -1 ? rA : rB, jid);
}

// b2MergeIsland, S/island.c:446-531: relabel the child's bodies, contacts and
// joints and append its lists to the root's. relabel false: the caller has
// relabelled them already (the labels do not take part in the appends).
F2D_HD void mergeIsland(IslandCtx& c, int id, bool relabel = true) {
  Island& is = c.islands[id];
  int rootId = is.parent;
  Island& root = c.islands[rootId];
  if (relabel) {
    int guard = c.h->bodyCount;
    for (int b = is.headBody; b != -1 && guard-- > 0; b = c.sims[b].islandNext) c.sims[b].islandId = rootId;
    guard = c.h->contactCap;
    for (int k = is.headContact; k != -1 && guard-- > 0; k = c.links[k].z) c.links[k].x = rootId;
    guard = c.h->jointCount;
    for (int j = is.headJoint; j != -1 && guard-- > 0; j = c.joints[j].islandNext) c.joints[j].islandId = rootId;
  }
  c.sims[root.tailBody].islandNext = is.headBody;
  c.sims[is.headBody].islandPrev = root.tailBody;
  root.tailBody = is.tailBody;
  root.bodyCount += is.bodyCount;
  if (root.headContact == -1) {
    root.headContact = is.headContact;
    root.tailContact = is.tailContact;
    root.contactCount = is.contactCount;
  } else if (is.headContact != -1) {
    c.links[root.tailContact].z = is.headContact;
    c.links[is.headContact].y = root.tailContact;
    root.tailContact = is.tailContact;
    root.contactCount += is.contactCount;
  }
  if (root.headJoint == -1) {
    root.headJoint = is.headJoint;
    root.tailJoint = is.tailJoint;
    root.jointCount = is.jointCount;
  } else if (is.headJoint != -1) {
    c.joints[root.tailJoint].islandNext = is.headJoint;
    c.joints[is.headJoint].islandPrev = root.tailJoint;
    root.tailJoint = is.tailJoint;
    root.jointCount += is.jointCount;
  }
  root.removeCount += is.removeCount;
}

// b2MergeAwakeIslands, S/island.c:537-592
F2D_HD_COLD void mergeAwakeIslands(IslandCtx& c) {
  WorldHeader* h = c.h;
  // without a parent link since the last merge every awake island is a root and the
  // merge changes nothing
  if (!h->mergePending) return;
  h->mergePending = 0;
  for (int i = 0; i < h->awakeIslandCount; ++i) {
    int id = c.awakeIslands[i];
    int root = islandRoot(c, id);
    if (root != id) c.islands[id].parent = root;
  }
  for (int i = h->awakeIslandCount - 1; i >= 0; --i) {
    int id = c.awakeIslands[i];
    if (c.islands[id].parent == -1) continue;
    mergeIsland(c, id);
    destroyIsland(c, id);
  }
}

// The merges of mergeAwakeIslands once every awake island points at its root and
// the children's bodies, contacts and joints carry their roots' ids (the solve
// kernel does both on all lanes): children merge into their roots in reverse
// awake order.
F2D_HD_COLD void mergeAwakeChildren(IslandCtx& c) {
  WorldHeader* h = c.h;
  h->mergePending = 0;
  for (int i = h->awakeIslandCount - 1; i >= 0; --i) {
    int id = c.awakeIslands[i];
    if (c.islands[id].parent == -1) continue;
    mergeIsland(c, id, false);
    destroyIsland(c, id);
  }
}

// b2SplitIsland, S/island.c:596-834: one depth-first search per unvisited body of
// the base island, following touching contacts (not into static bodies) and
// joints (into awake bodies). Marks live in bit 30 of a body's / contact's flags
// and in a joint's setIndex high bit; they are cleared for the base island first.
constexpr int kMarkBit = 1 << 30;

F2D_HD_COLD void splitIsland(IslandCtx& c, int baseId) {
  Island& base = c.islands[baseId];
  if (base.setIndex != kAwakeSet || base.removeCount == 0) return;
  const int n = base.bodyCount;
  int k = 0;
  for (int b = base.headBody; b != -1 && k < n; b = c.sims[b].islandNext) {
    c.seeds[k++] = b;
    c.sims[b].flags &= ~kMarkBit;
  }
  int guard = c.h->contactCap;
  for (int q = base.headContact; q != -1 && guard-- > 0; q = c.links[q].z) c.contacts[q].flags &= ~kMarkBit;
  guard = c.h->jointCount;
  for (int j = base.headJoint; j != -1 && guard-- > 0; j = c.joints[j].islandNext) c.joints[j].setIndex &= ~kMarkBit;
  destroyIsland(c, baseId);
  for (int i = 0; i < k; ++i) {
    int seed = c.seeds[i];
    if (c.sims[seed].flags & kMarkBit) continue;
    int top = 0;
    c.stack[top++] = seed;
    c.sims[seed].flags |= kMarkBit;
    int id = createIsland(c);
    if (id < 0) break;  // cannot happen: the split frees the base island first
    Island& is = c.islands[id];
    while (top > 0) {
      int bid = c.stack[--top];
      BodySim& body = c.sims[bid];
      body.islandId = id;
      if (is.tailBody != -1) c.sims[is.tailBody].islandNext = bid;
      body.islandPrev = is.tailBody;
      body.islandNext = -1;
      is.tailBody = bid;
      if (is.headBody == -1) is.headBody = bid;
      is.bodyCount += 1;
      for (int key = body.headContactKey, g = 2 * c.h->contactCap; key != -1 && g > 0; --g) {
        int cid = key >> 1, edge = key & 1;
        Contact& ct = c.contacts[cid];
        key = edge ? ct.nextB : ct.nextA;
        if (ct.flags & kMarkBit) continue;
        if ((ct.flags & CF_TOUCHING) == 0) continue;
        ct.flags |= kMarkBit;
        int other = edge ? ct.bodyA : ct.bodyB;
        BodySim& ob = c.sims[other];
        if ((ob.flags & kMarkBit) == 0 && ob.setIndex != kStaticSet && top < n) {
          c.stack[top++] = other;
          ob.flags |= kMarkBit;
        }
        int4& l = c.links[cid];
        l.x = id;
        if (is.tailContact != -1) c.links[is.tailContact].z = cid;
        l.y = is.tailContact;
        l.z = -1;
        is.tailContact = cid;
        if (is.headContact == -1) is.headContact = cid;
        is.contactCount += 1;
      }
      for (int key = body.headJointKey, g = 2 * c.h->jointCount; key != -1 && g > 0; --g) {
        int jid = key >> 1, edge = key & 1;
        Joint& j = c.joints[jid];
        key = edge ? j.nextB : j.nextA;
        if (j.setIndex & kMarkBit) continue;
        j.setIndex |= kMarkBit;
        int other = edge ? j.bodyA : j.bodyB;
        BodySim& ob = c.sims[other];
        if (ob.setIndex == kDisabledSet) continue;
        if ((ob.flags & kMarkBit) == 0 && ob.setIndex == kAwakeSet && top < n) {
          c.stack[top++] = other;
          ob.flags |= kMarkBit;
        }
        j.islandId = id;
        if (is.tailJoint != -1) c.joints[is.tailJoint].islandNext = jid;
        j.islandPrev = is.tailJoint;
        j.islandNext = -1;
        is.tailJoint = jid;
        if (is.headJoint == -1) is.headJoint = jid;
        is.jointCount += 1;
      }
    }
  }
  // clear the marks again (bodies of the base island, contacts and joints of the new islands)
  for (int i = 0; i < k; ++i) c.sims[c.seeds[i]].flags &= ~kMarkBit;
  for (int i = 0; i < k; ++i) {
    const BodySim& body = c.sims[c.seeds[i]];
    for (int key = body.headContactKey, g = 2 * c.h->contactCap; key != -1 && g > 0; --g) {
      Contact& ct = c.contacts[key >> 1];
      key = (key & 1) ? ct.nextB : ct.nextA;
      ct.flags &= ~kMarkBit;
    }
    for (int key = body.headJointKey, g = 2 * c.h->jointCount; key != -1 && g > 0; --g) {
      Joint& j = c.joints[key >> 1];
      key = (key & 1) ? j.nextB : j.nextA;
      j.setIndex &= ~kMarkBit;
    }
  }
}

}  // namespace f2d
