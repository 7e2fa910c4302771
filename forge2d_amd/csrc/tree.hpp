// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// Broad-phase bounding-volume trees of one world (b2DynamicTree, S/dynamic_tree.c).
//
// Each world keeps the reference's three trees (static, kinematic and dynamic
// proxies). The set of overlapping pairs does not depend on them -- the device
// finds it by sort and sweep -- but the reference numbers the new contacts of one
// moved proxy in the order its tree query visits the partners (depth first, child2
// before child1, S/dynamic_tree.c:1114-1170), so the trees are kept exactly:
// insertion with rotations and removal (host, when proxies are created or
// destroyed), enlargement of moved leaves and the partial rebuild every step
// (device), and each leaf's rank in that visiting order.
//
// The node free list is held as the freed node ids (a stack) above the never-used
// ids [fresh, capacity), which the reference's list holds in ascending order:
// popping the stack, then taking fresh ids, hands out the ids the reference does.
//
// These functions run on the host (proxy creation, snapshot import, the
// enlargements a device step left pending); the device's per-step rebuild is the
// block-parallel restatement in tree_device.hpp.
#pragma once

#include "b2math.hpp"

namespace f2d {

enum : int { TN_ALLOCATED = 1, TN_ENLARGED = 2, TN_LEAF = 4 };  // S/constants.h:55-60

// b2TreeNode, S/dynamic_tree.c:19-50. A leaf's user data (the shape id) is child1.
// Category bits are not kept: an internal node's are the union of its leaves', and
// the pair queries use the default mask, which no non-zero category misses.
struct alignas(16) TreeNode {
  AABB box;
  int child1, child2, parent;
  uint16_t height, flags;
};
static_assert(sizeof(TreeNode) == 32, "tree node");

struct alignas(16) TreeHeader {
  int root;
  int nodeCount;   // allocated nodes
  int proxyCount;
  int fresh;       // next never-used node id
  int freedCount;  // depth of the freed-id stack
  int cap;         // node capacity of this tree's region
  int offset;      // region start in the batch's node and freed-id arrays
  int error;       // set when the region's capacity was exceeded
};

// ints of the device rebuild's work arrays for F node slots and P proxies
F2D_HD long treeWorkInts(int F, int P) { return 4L * F + 20L * P + 160; }

struct TreeRef {
  TreeHeader* h;
  TreeNode* n;
  int* freed;
};

F2D_HD float perimeter(AABB a) {  // b2Perimeter, S/aabb.h
  float wx = a.hi.x - a.lo.x;
  float wy = a.hi.y - a.lo.y;
  return 2.0f * (wx + wy);
}
F2D_HD AABB aabbUnion(AABB a, AABB b) {  // b2AABB_Union, I/math_functions.h:607-615
  AABB c;
  c.lo.x = minf(a.lo.x, b.lo.x);
  c.lo.y = minf(a.lo.y, b.lo.y);
  c.hi.x = maxf(a.hi.x, b.hi.x);
  c.hi.y = maxf(a.hi.y, b.hi.y);
  return c;
}
F2D_HD Vec2 aabbCenter(AABB a) {  // b2AABB_Center :593-597
  Vec2 c = {0.5f * (a.lo.x + a.hi.x), 0.5f * (a.lo.y + a.hi.y)};
  return c;
}
F2D_HD bool aabbEnlarge(AABB* a, AABB b) {  // b2EnlargeAABB, S/aabb.h
  bool changed = false;
  if (b.lo.x < a->lo.x) { a->lo.x = b.lo.x; changed = true; }
  if (b.lo.y < a->lo.y) { a->lo.y = b.lo.y; changed = true; }
  if (a->hi.x < b.hi.x) { a->hi.x = b.hi.x; changed = true; }
  if (a->hi.y < b.hi.y) { a->hi.y = b.hi.y; changed = true; }
  return changed;
}

// b2AllocateNode, S/dynamic_tree.c:122-158 (the caller guarantees capacity; -1 and
// the header's error flag when it runs out)
F2D_HD int treeAlloc(TreeRef t) {
  int id;
  if (t.h->freedCount > 0) {
    id = t.freed[--t.h->freedCount];
  } else {
    if (t.h->fresh >= t.h->cap) {
      t.h->error = 1;
      return -1;
    }
    id = t.h->fresh++;
  }
  TreeNode& n = t.n[id];
  n.box = AABB{{0.0f, 0.0f}, {0.0f, 0.0f}};
  n.child1 = n.child2 = n.parent = -1;
  n.height = 0;
  n.flags = TN_ALLOCATED;
  t.h->nodeCount += 1;
  return id;
}

// b2FreeNode :161-170
F2D_HD void treeFree(TreeRef t, int id) {
  t.freed[t.h->freedCount++] = id;
  t.n[id].flags = 0;
  t.h->nodeCount -= 1;
}

// b2FindBestSibling :186-330: one greedy descent by surface-area cost
F2D_HD int findBestSibling(const TreeNode* n, int root, AABB boxD) {
  const Vec2 centerD = aabbCenter(boxD);
  const float areaD = perimeter(boxD);
  int index = root;
  float areaBase = perimeter(n[index].box);
  float directCost = perimeter(aabbUnion(n[index].box, boxD));
  float inherited = 0.0f;
  int best = index;
  float bestCost = directCost;
  while (n[index].height > 0) {
    const int c1 = n[index].child1, c2 = n[index].child2;
    float cost = directCost + inherited;
    if (cost < bestCost) {
      best = index;
      bestCost = cost;
    }
    inherited += directCost - areaBase;
    const bool leaf1 = n[c1].height == 0, leaf2 = n[c2].height == 0;
    const AABB b1 = n[c1].box, b2 = n[c2].box;
    float lower1 = kFltMax, lower2 = kFltMax, area1 = 0.0f, area2 = 0.0f;
    const float direct1 = perimeter(aabbUnion(b1, boxD));
    if (leaf1) {
      float cost1 = direct1 + inherited;
      if (cost1 < bestCost) {
        best = c1;
        bestCost = cost1;
      }
    } else {
      area1 = perimeter(b1);
      lower1 = inherited + direct1 + minf(areaD - area1, 0.0f);
    }
    const float direct2 = perimeter(aabbUnion(b2, boxD));
    if (leaf2) {
      float cost2 = direct2 + inherited;
      if (cost2 < bestCost) {
        best = c2;
        bestCost = cost2;
      }
    } else {
      area2 = perimeter(b2);
      lower2 = inherited + direct2 + minf(areaD - area2, 0.0f);
    }
    if (leaf1 && leaf2) break;
    if (bestCost <= lower1 && bestCost <= lower2) break;
    if (lower1 == lower2 && !leaf1) {
      // both children contain D as cheaply: the one whose centre is nearer
      lower1 = lengthSq(sub(aabbCenter(b1), centerD));
      lower2 = lengthSq(sub(aabbCenter(b2), centerD));
    }
    if (lower1 < lower2 && !leaf1) {
      index = c1;
      areaBase = area1;
      directCost = direct1;
    } else {
      index = c2;
      areaBase = area2;
      directCost = direct2;
    }
  }
  return best;
}

// One child swap of b2RotateNodes: `out`, a child of A, trades places with `in`, a
// child of A's other child P whose remaining child is `keep`; P takes `box`.
F2D_HD void treeSwap(TreeNode* n, int a, int p, int out, int in, int keep, AABB box, bool outFirst,
                     bool inFirst) {
  if (outFirst) n[a].child1 = in; else n[a].child2 = in;
  if (inFirst) n[p].child1 = out; else n[p].child2 = out;
  n[out].parent = p;
  n[in].parent = a;
  n[p].box = box;
  n[p].height = (uint16_t)(1 + maxi(n[out].height, n[keep].height));
  n[a].height = (uint16_t)(1 + maxi(n[p].height, n[in].height));
  n[p].flags |= (n[out].flags | n[keep].flags) & TN_ENLARGED;
  n[a].flags |= (n[p].flags | n[in].flags) & TN_ENLARGED;
}

// b2RotateNodes :340-594
F2D_HD void rotateNodes(TreeNode* n, int a) {
  if (n[a].height < 2) return;
  const int b = n[a].child1, c = n[a].child2;
  if (n[b].height == 0) {
    const int f = n[c].child1, g = n[c].child2;
    const float base = perimeter(n[c].box);
    const AABB bg = aabbUnion(n[b].box, n[g].box), bf = aabbUnion(n[b].box, n[f].box);
    const float costBF = perimeter(bg), costBG = perimeter(bf);
    if (base < costBF && base < costBG) return;
    if (costBF < costBG) treeSwap(n, a, c, b, f, g, bg, true, true);
    else treeSwap(n, a, c, b, g, f, bf, true, false);
  } else if (n[c].height == 0) {
    const int d = n[b].child1, e = n[b].child2;
    const float base = perimeter(n[b].box);
    const AABB ce = aabbUnion(n[c].box, n[e].box), cd = aabbUnion(n[c].box, n[d].box);
    const float costCD = perimeter(ce), costCE = perimeter(cd);
    if (base < costCD && base < costCE) return;
    if (costCD < costCE) treeSwap(n, a, b, c, d, e, ce, false, true);
    else treeSwap(n, a, b, c, e, d, cd, false, false);
  } else {
    const int d = n[b].child1, e = n[b].child2, f = n[c].child1, g = n[c].child2;
    const float areaB = perimeter(n[b].box), areaC = perimeter(n[c].box);
    float bestCost = areaB + areaC;
    int best = 0;
    const AABB bg = aabbUnion(n[b].box, n[g].box);
    const float costBF = areaB + perimeter(bg);
    if (costBF < bestCost) { best = 1; bestCost = costBF; }
    const AABB bf = aabbUnion(n[b].box, n[f].box);
    const float costBG = areaB + perimeter(bf);
    if (costBG < bestCost) { best = 2; bestCost = costBG; }
    const AABB ce = aabbUnion(n[c].box, n[e].box);
    const float costCD = areaC + perimeter(ce);
    if (costCD < bestCost) { best = 3; bestCost = costCD; }
    const AABB cd = aabbUnion(n[c].box, n[d].box);
    const float costCE = areaC + perimeter(cd);
    if (costCE < bestCost) best = 4;
    if (best == 1) treeSwap(n, a, c, b, f, g, bg, true, true);
    else if (best == 2) treeSwap(n, a, c, b, g, f, bf, true, false);
    else if (best == 3) treeSwap(n, a, b, c, d, e, ce, false, true);
    else if (best == 4) treeSwap(n, a, b, c, e, d, cd, false, false);
  }
}

F2D_HD void refitNode(TreeNode* n, int i) {
  const int c1 = n[i].child1, c2 = n[i].child2;
  n[i].box = aabbUnion(n[c1].box, n[c2].box);
  n[i].height = (uint16_t)(1 + maxi(n[c1].height, n[c2].height));
}

// b2InsertLeaf :625-701
F2D_HD bool insertLeaf(TreeRef t, int leaf, bool rotate) {
  TreeNode* n = t.n;
  if (t.h->root < 0) {
    t.h->root = leaf;
    n[leaf].parent = -1;
    return true;
  }
  const AABB leafBox = n[leaf].box;
  const int sibling = findBestSibling(n, t.h->root, leafBox);
  const int oldParent = n[sibling].parent;
  const int np = treeAlloc(t);
  if (np < 0) return false;
  n[np].parent = oldParent;
  n[np].box = aabbUnion(leafBox, n[sibling].box);
  n[np].height = (uint16_t)(n[sibling].height + 1);
  if (oldParent >= 0) {
    if (n[oldParent].child1 == sibling) n[oldParent].child1 = np; else n[oldParent].child2 = np;
  } else {
    t.h->root = np;
  }
  n[np].child1 = sibling;
  n[np].child2 = leaf;
  n[sibling].parent = np;
  n[leaf].parent = np;
  for (int i = np; i >= 0; i = n[i].parent) {
    refitNode(n, i);
    n[i].flags |= (n[n[i].child1].flags | n[n[i].child2].flags) & TN_ENLARGED;
    if (rotate) rotateNodes(n, i);
  }
  return true;
}

// b2RemoveLeaf :703-768
F2D_HD void removeLeaf(TreeRef t, int leaf) {
  TreeNode* n = t.n;
  if (leaf == t.h->root) {
    t.h->root = -1;
    return;
  }
  const int parent = n[leaf].parent;
  const int grand = n[parent].parent;
  const int sibling = n[parent].child1 == leaf ? n[parent].child2 : n[parent].child1;
  if (grand >= 0) {
    if (n[grand].child1 == parent) n[grand].child1 = sibling; else n[grand].child2 = sibling;
    n[sibling].parent = grand;
    treeFree(t, parent);
    for (int i = grand; i >= 0; i = n[i].parent) refitNode(n, i);
  } else {
    t.h->root = sibling;
    n[sibling].parent = -1;
    treeFree(t, parent);
  }
}

// b2DynamicTree_CreateProxy :772-795; returns the proxy id, -1 without capacity
F2D_HD int treeCreateProxy(TreeRef t, AABB box, int shapeId) {
  const int id = treeAlloc(t);
  if (id < 0) return -1;
  TreeNode& n = t.n[id];
  n.box = box;
  n.child1 = shapeId;
  n.child2 = 0;
  n.height = 0;
  n.flags = TN_ALLOCATED | TN_LEAF;
  if (!insertLeaf(t, id, true)) return -1;
  t.h->proxyCount += 1;
  return id;
}

// b2DynamicTree_DestroyProxy :797-809
F2D_HD void treeDestroyProxy(TreeRef t, int id) {
  removeLeaf(t, id);
  treeFree(t, id);
  t.h->proxyCount -= 1;
}

// b2DynamicTree_MoveProxy :811-825 (no rotations)
F2D_HD bool treeMoveProxy(TreeRef t, int id, AABB box) {
  removeLeaf(t, id);
  t.n[id].box = box;
  return insertLeaf(t, id, false);
}

// b2DynamicTree_EnlargeProxy :827-866
F2D_HD void treeEnlargeProxy(TreeRef t, int id, AABB box) {
  TreeNode* n = t.n;
  n[id].box = box;
  int p = n[id].parent;
  while (p >= 0) {
    const bool changed = aabbEnlarge(&n[p].box, box);
    n[p].flags |= TN_ENLARGED;
    p = n[p].parent;
    if (!changed) break;
  }
  while (p >= 0 && (n[p].flags & TN_ENLARGED) == 0) {
    n[p].flags |= TN_ENLARGED;
    p = n[p].parent;
  }
}

// Each leaf's position in the order a query visits leaves (depth first, child2
// before child1): rank[shape id] for every proxy of the tree. `walk` holds the
// tree's height plus one.
F2D_HD void treeLeafRanks(const TreeNode* n, int root, int* walk, int* rankOfShape) {
  if (root < 0) return;
  int sp = 0, r = 0;
  walk[sp++] = root;
  while (sp > 0) {
    const int i = walk[--sp];
    if (n[i].flags & TN_LEAF) {
      rankOfShape[n[i].child1] = r++;
    } else {
      walk[sp++] = n[i].child1;
      walk[sp++] = n[i].child2;
    }
  }
}

}  // namespace f2d
