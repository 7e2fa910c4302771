// Restates algorithms of Box2D v3.1.1, MIT License, Copyright (c) 2022 Erin Catto (THIRD_PARTY_NOTICES).
// The partial tree rebuild of one world on one wavefront (b2DynamicTree_Rebuild with
// fullBuild false, S/dynamic_tree.c:1717-1989), with the result of the reference's
// serial code: the same nodes, links, boxes, heights, node ids and free list.
//
// The serial rebuild collects, in child1-first depth-first order, the leaves and
// the internal nodes that did not grow (the build's items), frees the grown
// internal nodes in the same order, and builds a subtree over the items by median
// splits (Hoare partitions of the item centres), allocating a node per range of two
// or more items in depth-first order. Here every step is a pass of the wavefront's
// lanes:
//
// * Items. The grown internal nodes form the top of the tree; an item is a node
//   outside them whose parent is inside. The leaves of an item's subtree hold a
//   contiguous interval of the last ranks (the query order, child2 first), whose
//   upper end is reached by descending child1 and lower end by descending child2:
//   ordering items by that upper end, descending, is the child1-first order.
// * Freed order. A grown node's place in child1-first preorder among the grown
//   nodes is the place of its first item plus the number of child1 links on its
//   path from the root. The build pops them in reverse (the free list is LIFO),
//   so the build's k-th node is the grown node at place (grown - 1 - k).
// * Build. All ranges of one depth are partitioned at once. The Hoare scheme of
//   b2PartitionMid swaps the k-th item at or above the pivot in the left part with
//   the k-th item below it in the right part, counted from the right end; the
//   ranks of both come from one prefix count of the "below the pivot" flags over
//   all items. A range's node is the build's node at preorder place (range start +
//   number of child1 links above it).
// * Ranks. The query visits the items from the last to the first, each item's
//   leaves in their old relative order: a leaf's new rank is the number of leaves
//   of later items plus its old rank's offset within its item's interval.
#pragma once

#include "layout.hpp"

namespace f2d {

// floats as ints with the order of `<` (negative zero below positive zero: min and
// max of centre bounds may differ from the serial scan only in the sign of a zero,
// which neither the axis test nor the pivot comparisons can see)
__device__ __forceinline__ int floatOrderKey(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float orderKeyFloat(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

// Work arrays of one tree rebuild for F node slots and P proxies (ints).
struct TreeWork {
  int4* links;     // [F] child1, child2, parent, flags | item place << 8
  int* items;      // [P] build items, permuted by the partitions
  float* cx;       // [P] item centres (cx, cy, seg and pre are contiguous: 16 B per item,
  float* cy;       //     reused as range boxes by the refit)
  int* seg;        // [P] range of each item place at the current depth, -1 when done
  int* pre;        // [P + 64] prefix counts of the pivot flags
  int* partner;    // [P] right-part places by rank; heights in the refit; final places in the ranks
  int4* ranges;    // [P] start, end, child1 links above, node id
  int* split;      // [P] split place
  int* left;       // [P] child1 range index, or -1 - item node
  int* right;      // [P] child2 range index, or -1 - item node
  int* bounds;     // [2P + 8] per range of a depth: centre bounds as order keys, then axis and
                   //     pivot (aliases byRank and lowOf, idle during the build)
  int* grown;      // [P] grown nodes in child1-first preorder
  int* byRank;     // [P] item by the upper end of its rank interval, then each rank's owner
  int* lowOf;      // [P] lower end by upper end, then leaves after each item
  int* leaves;     // [P] leaves of the item at each place
  int* lowMark;    // [P] item place by the lower end of its rank interval
  int* levels;     // [P + 2] first range of each depth
  float4* itemBox; // [P] (aliases grown, byRank, lowOf, split) item boxes by first place, for the refit
  int* itemHeight; // [P] item heights by first place
  int* counter;    // [4] range count
  int* sums;       // [16] per-wavefront partial sums of the block scans
};

__device__ __forceinline__ TreeWork treeWorkAt(int* base, int F, int P) {
  TreeWork t;
  int* p = base;
  t.links = reinterpret_cast<int4*>(p);
  p += 4L * F;
  t.cx = reinterpret_cast<float*>(p);
  p += P;
  t.cy = reinterpret_cast<float*>(p);
  p += P;
  t.seg = p;
  p += P;
  t.pre = p;
  p += P + 64;
  t.ranges = reinterpret_cast<int4*>(p);
  p += 4L * P;
  t.items = p;
  p += P;
  t.partner = p;
  p += P;
  t.left = p;
  p += P;
  t.right = p;
  p += P;
  t.grown = p;
  p += P;
  t.byRank = p;
  p += P;
  t.lowOf = p;
  p += P + 8;
  t.bounds = t.byRank;  // the build runs between the item and rank phases that use them
  t.split = p;
  p += P;
  // item boxes for the refit, by first place, over grown..split (idle after the build)
  t.itemBox = reinterpret_cast<float4*>(t.grown);
  t.itemHeight = p;
  p += P;
  t.leaves = p;
  p += P;
  t.lowMark = p;
  p += P;
  t.levels = p;
  p += P + 2;
  t.counter = p;
  p += 4;
  t.sums = p;
  return t;
}

__device__ __forceinline__ unsigned long long laneMaskLt(int lane) { return (1ull << lane) - 1ull; }

// inclusive scans across the 64 lanes of a wavefront
__device__ __forceinline__ int waveInclusiveAdd(int v, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}
__device__ __forceinline__ int waveInclusiveMax(int v, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    if (lane >= d) v = max(v, o);
  }
  return v;
}

// Minimum (or maximum) over the run of lanes that share `key` with this lane and
// precede it: lanes of one key are consecutive (items of one range are consecutive
// places), so the run's last lane holds the run's result.
__device__ __forceinline__ int segMinUp(int v, int key, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    const int k = __shfl_up(key, d, 64);
    if (lane >= d && k == key) v = min(v, o);
  }
  return v;
}
__device__ __forceinline__ int segMaxUp(int v, int key, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    const int k = __shfl_up(key, d, 64);
    if (lane >= d && k == key) v = max(v, o);
  }
  return v;
}

// Block-wide scans over BLOCK lanes (BLOCK / 64 wavefronts): every lane of the
// workgroup calls them, in the same order. `sums` holds one partial per wavefront.
template <int BLOCK>
__device__ __forceinline__ int blockFlagPrefix(bool f, int* sums, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(f);
  const int prefix = __popcll(m & laneMaskLt(lane));
  if (BLOCK == 64) {
    *total = __popcll(m);
    return prefix;
  }
  if (lane == 0) sums[wave] = __popcll(m);
  __syncthreads();
  int base = 0, sum = 0;
  for (int k = 0; k < BLOCK / 64; ++k) {
    const int v = sums[k];
    base += k < wave ? v : 0;
    sum += v;
  }
  __syncthreads();
  *total = sum;
  return base + prefix;
}
template <int BLOCK>
__device__ __forceinline__ int blockInclusiveAdd(int v, int* sums, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = waveInclusiveAdd(v, lane);
  if (BLOCK == 64) {
    *total = __shfl(incl, 63, 64);
    return incl;
  }
  if (lane == 63) sums[wave] = incl;
  __syncthreads();
  int base = 0, sum = 0;
  for (int k = 0; k < BLOCK / 64; ++k) {
    const int x = sums[k];
    base += k < wave ? x : 0;
    sum += x;
  }
  __syncthreads();
  *total = sum;
  return base + incl;
}
template <int BLOCK>
__device__ __forceinline__ int blockInclusiveMax(int v, int* sums, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = waveInclusiveMax(v, lane);
  if (BLOCK == 64) {
    *total = __shfl(incl, 63, 64);
    return incl;
  }
  if (lane == 63) sums[wave] = incl;
  __syncthreads();
  int base = -1, all = -1;
  for (int k = 0; k < BLOCK / 64; ++k) {
    const int x = sums[k];
    base = k < wave ? max(base, x) : base;
    all = max(all, x);
  }
  __syncthreads();
  *total = all;
  return max(base, incl);
}

// One tree of one world: the leaves the last step enlarged take their shapes' fat
// boxes and mark their ancestors grown, then the partial rebuild and the new ranks.
// gn: the tree's nodes in HBM; ranks: the world's per-shape rank array. Called by
// all BLOCK lanes of the workgroup.
template <int BLOCK>
__device__ __forceinline__ void rebuildTree(TreeHeader* hd, TreeNode* gn, Shape* shapes, int shapeCount,
                                            int treeType, int* ranks, TreeWork w) {
  const int lane = threadIdx.x;
  const int F = hd->fresh, P = hd->proxyCount, root = hd->root;
  if (P == 0 || root < 0 || root >= F) return;
  // a tree that disagrees with its ranks (never expected) is left unbuilt with the
  // header's error flag set; the flag is agreed on by a wavefront vote
  bool bad = false;
  for (int i = lane; i < F; i += BLOCK) {
    const TreeNode& n = gn[i];
    w.links[i] = make_int4(n.child1, n.child2, n.parent, (int)n.flags);
  }
  __syncthreads();
  // enlarged leaves (b2DynamicTree_EnlargeProxy, S/dynamic_tree.c:827-866): a walk up
  // stops at a node another walk marked, whose ancestors that walk marks
  bool any = false;
  for (int k = lane; k < shapeCount; k += BLOCK) {
    Shape& sh = shapes[k];
    const int flags = sh.flags;
    if (!(flags & SF_TREE_PENDING) || sh.proxyKey < 0 || (sh.proxyKey & 3) != treeType) continue;
    sh.flags = flags & ~SF_TREE_PENDING;
    const int leaf = sh.proxyKey >> 2;
    if (leaf >= F) {
      bad = true;
      continue;
    }
    gn[leaf].box = sh.fat;
    any = true;
    for (int p = w.links[leaf].z, g = 0; p >= 0 && p < F && g < F; p = w.links[p].z, ++g) {
      const int old = atomicOr(&w.links[p].w, TN_ENLARGED);
      if (old & TN_ENLARGED) break;
    }
  }
  (void)any;
  __syncthreads();
  if (__syncthreads_or(bad) || hd->error) {
    if (lane == 0) hd->error = 2;
    return;
  }
  const int4 rootLink = w.links[root];
  if ((rootLink.w & TN_LEAF) || !(rootLink.w & TN_ENLARGED)) return;

  // ---- items and their rank intervals
  for (int r = lane; r < P; r += BLOCK) {
    w.byRank[r] = -1;
    w.lowMark[r] = -1;
  }
  __syncthreads();
  for (int i = lane; i < F; i += BLOCK) {
    const int4 n = w.links[i];
    if (!(n.w & TN_ALLOCATED)) continue;
    if (!(n.w & TN_LEAF) && (n.w & TN_ENLARGED)) continue;
    if (n.z < 0 || n.z >= F || !(w.links[n.z].w & TN_ENLARGED)) continue;
    // walks stay inside the node range and are bounded by the node count
    int a = i, b = i;
    for (int g = 0; g < F && !(w.links[a].w & TN_LEAF); ++g) {
      const int c = w.links[a].x;
      if (c < 0 || c >= F) break;
      a = c;
    }
    for (int g = 0; g < F && !(w.links[b].w & TN_LEAF); ++g) {
      const int c = w.links[b].y;
      if (c < 0 || c >= F) break;
      b = c;
    }
    if (!(w.links[a].w & TN_LEAF) || !(w.links[b].w & TN_LEAF)) {
      bad = true;
      continue;
    }
    const int sa = w.links[a].x, sb = w.links[b].x;
    if (sa < 0 || sa >= shapeCount || sb < 0 || sb >= shapeCount) {
      bad = true;
      continue;
    }
    const int hi = ranks[sa], lo = ranks[sb];
    if (hi < 0 || hi >= P || lo < 0 || lo > hi) {
      bad = true;
      continue;
    }
    w.byRank[hi] = i;
    w.lowOf[hi] = lo;
  }
  __syncthreads();
  int count = 0;
  for (int base = P - 1; base >= 0; base -= BLOCK) {
    const int r = base - lane;
    const bool v = r >= 0 && w.byRank[r] >= 0;
    int total;
    const int prefix = blockFlagPrefix<BLOCK>(v, w.sums, &total);
    if (v) {
      const int pos = count + prefix;
      const int node = w.byRank[r];
      const int lo = w.lowOf[r];
      w.items[pos] = node;
      w.leaves[pos] = r - lo + 1;
      w.lowMark[lo] = pos;
      w.links[node].w = (w.links[node].w & 0xFF) | (pos << 8);
      const AABB bx = gn[node].box;
      w.cx[pos] = 0.5f * (bx.lo.x + bx.hi.x);
      w.cy[pos] = 0.5f * (bx.lo.y + bx.hi.y);
      w.seg[pos] = 0;
    }
    count += total;
  }
  __syncthreads();
  if (__syncthreads_or(bad) || count < 2) {
    if (lane == 0) hd->error = 2;
    return;
  }
  const int nGrown = count - 1;
  for (int k = lane; k < nGrown; k += BLOCK) w.grown[k] = -1;
  __syncthreads();

  // ---- grown nodes in child1-first preorder
  for (int i = lane; i < F; i += BLOCK) {
    const int4 n = w.links[i];
    if (!(n.w & TN_ALLOCATED) || (n.w & TN_LEAF) || !(n.w & TN_ENLARGED)) continue;
    int a = n.x;
    for (int g = 0; g < F && a >= 0 && a < F && !(w.links[a].w & TN_LEAF) && (w.links[a].w & TN_ENLARGED); ++g)
      a = w.links[a].x;
    int turns = 0;
    for (int c = i, p = n.z, g = 0; p >= 0 && p < F && g < F; c = p, p = w.links[p].z, ++g)
      turns += w.links[p].x == c ? 1 : 0;
    const int place = (a >= 0 && a < F) ? (w.links[a].w >> 8) + turns : -1;
    if (place >= 0 && place < nGrown) {
      w.grown[place] = i;
    } else {
      bad = true;
    }
  }
  if (lane == 0) {
    w.ranges[0] = make_int4(0, count, 0, 0);
    w.levels[0] = 0;
    w.levels[1] = 1;
    w.counter[0] = 1;
  }
  __syncthreads();
  // every place holds a grown node, each a node of this tree: otherwise the tree and
  // the ranks disagree, and the tree is left as it is (error flag)
  for (int k = lane; k < nGrown; k += BLOCK) {
    const int g = w.grown[k];
    if (g < 0 || g >= F) bad = true;
  }
  if (__syncthreads_or(bad)) {
    if (lane == 0) hd->error = 2;
    return;
  }
  if (lane == 0) w.ranges[0].w = w.grown[nGrown - 1];

  // ---- the build, one depth at a time
  int begin = 0, end = 1, depth = 0, prevBegin = 0;
  if (lane == 0) {
    w.bounds[0] = w.bounds[1] = 0x7FFFFFFF;
    w.bounds[2] = w.bounds[3] = (int)0x80000000;
  }
  __syncthreads();
  while (begin < end && depth < count) {
    const int nl = end - begin;
    // each item's range at this depth (the child of its last range), and the centre
    // bounds of every range of three or more items
    for (int base = 0; base < count; base += BLOCK) {
      const int i = base + lane;
      int r = -1;
      if (i < count) {
        r = w.seg[i];
        if (depth > 0) {
          if (r >= prevBegin) {
            const int c = i < w.split[r] ? w.left[r] : w.right[r];
            r = c >= 0 ? c : -1;
            w.seg[i] = r;
          } else {
            r = -1;
          }
        }
        if (r >= 0) {
          const int4 R = w.ranges[r];
          if (R.y - R.x < 3) r = -1;
        }
      }
      // one atomic per range and wavefront: reduce each run of lanes of one range first
      const int wl = lane & 63;
      const int kx = r >= 0 ? floatOrderKey(w.cx[i]) : 0, ky = r >= 0 ? floatOrderKey(w.cy[i]) : 0;
      const int mnx = segMinUp(kx, r, wl), mny = segMinUp(ky, r, wl);
      const int mxx = segMaxUp(kx, r, wl), mxy = segMaxUp(ky, r, wl);
      const int nextKey = __shfl_down(r, 1, 64);
      if (r >= 0 && (wl == 63 || nextKey != r)) {
        int* bd = w.bounds + 4 * (r - begin);
        atomicMin(bd + 0, mnx);
        atomicMin(bd + 1, mny);
        atomicMax(bd + 2, mxx);
        atomicMax(bd + 3, mxy);
      }
    }
    __syncthreads();
    // the longer axis of the centres' bounds and its middle (b2PartitionMid :1436-1450)
    for (int k = lane; k < nl; k += BLOCK) {
      const int4 R = w.ranges[begin + k];
      if (R.y - R.x < 3) continue;
      int* bd = w.bounds + 4 * k;
      const float lox = orderKeyFloat(bd[0]), loy = orderKeyFloat(bd[1]);
      const float hix = orderKeyFloat(bd[2]), hiy = orderKeyFloat(bd[3]);
      const bool useX = hix - lox > hiy - loy;
      const float pivot = useX ? 0.5f * (lox + hix) : 0.5f * (loy + hiy);
      bd[0] = useX ? 1 : 0;
      bd[1] = __float_as_int(pivot);
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < count; base += BLOCK) {
      const int i = base + lane;
      bool f = false;
      if (i < count) {
        const int r = w.seg[i];
        if (r >= begin) {
          const int4 R = w.ranges[r];
          if (R.y - R.x >= 3) {
            const int* bd = w.bounds + 4 * (r - begin);
            f = (bd[0] ? w.cx[i] : w.cy[i]) < __int_as_float(bd[1]);
          }
        }
      }
      int total;
      const int prefix = blockFlagPrefix<BLOCK>(f, w.sums, &total);
      if (i < count) w.pre[i] = carry + prefix;
      carry += total;
    }
    if (lane == 0) w.pre[count] = carry;
    __syncthreads();
    // a range of three or more items with items on both sides of the pivot splits
    // where the items below it end; the Hoare swaps pair the k-th item at or above
    // the pivot in the left part with the k-th below it in the right part, counted
    // from the right end
    for (int i = lane; i < count; i += BLOCK) {
      const int r = w.seg[i];
      if (r < begin) continue;
      const int4 R = w.ranges[r];
      const int below = w.pre[R.y] - w.pre[R.x];
      if (R.y - R.x < 3 || below == 0 || below == R.y - R.x) continue;
      if (i >= R.x + below && w.pre[i + 1] != w.pre[i]) w.partner[R.x + (w.pre[R.y] - w.pre[i + 1])] = i;
    }
    __syncthreads();
    for (int i = lane; i < count; i += BLOCK) {
      const int r = w.seg[i];
      if (r < begin) continue;
      const int4 R = w.ranges[r];
      const int below = w.pre[R.y] - w.pre[R.x];
      if (R.y - R.x < 3 || below == 0 || below == R.y - R.x) continue;
      if (i < R.x + below && w.pre[i + 1] == w.pre[i]) {
        const int j = w.partner[R.x + (i - R.x) - (w.pre[i] - w.pre[R.x])];
        const int ti = w.items[i];
        const float tx = w.cx[i], ty = w.cy[i];
        w.items[i] = w.items[j];
        w.cx[i] = w.cx[j];
        w.cy[i] = w.cy[j];
        w.items[j] = ti;
        w.cx[j] = tx;
        w.cy[j] = ty;
      }
    }
    __syncthreads();
    // the split; child ranges of two or more items get the build's node at their
    // preorder place and fresh bound slots for the next depth
    int next = end;
    for (int kb = 0; kb < nl; kb += BLOCK) {
      const int k = kb + lane;
      const int r = begin + k;
      int4 R = make_int4(0, 0, 0, 0);
      int m = 0, made = 0;
      if (k < nl) {
        R = w.ranges[r];
        const int size = R.y - R.x;
        m = R.x + size / 2;
        if (size >= 3) {
          const int below = w.pre[R.y] - w.pre[R.x];
          if (below > 0 && below < size) m = R.x + below;
        }
        made = (m - R.x >= 2 ? 1 : 0) + (R.y - m >= 2 ? 1 : 0);
      }
      // child ranges are numbered in range order (a block prefix of the children made)
      int total;
      const int slot = next + blockInclusiveAdd<BLOCK>(made, w.sums, &total) - made;
      next += total;
      if (k >= nl) continue;
      w.split[r] = m;
      int lc = -1 - w.items[R.x], rc = -1 - w.items[m];
      int nextSlot = slot;
      if (m - R.x >= 2) {
        lc = nextSlot++;
        const int k1 = mini(maxi(nGrown - 1 - (R.x + R.z + 1), 0), nGrown - 1);
        w.ranges[lc] = make_int4(R.x, m, R.z + 1, w.grown[k1]);
        int* bd = w.bounds + 4 * (lc - end);
        bd[0] = bd[1] = 0x7FFFFFFF;
        bd[2] = bd[3] = (int)0x80000000;
      }
      if (R.y - m >= 2) {
        rc = nextSlot++;
        const int k2 = mini(maxi(nGrown - 1 - (m + R.z), 0), nGrown - 1);
        w.ranges[rc] = make_int4(m, R.y, R.z, w.grown[k2]);
        int* bd = w.bounds + 4 * (rc - end);
        bd[0] = bd[1] = 0x7FFFFFFF;
        bd[2] = bd[3] = (int)0x80000000;
      }
      w.left[r] = lc;
      w.right[r] = rc;
    }
    __syncthreads();
    prevBegin = begin;
    begin = end;
    end = next;
    depth += 1;
    if (lane == 0) w.levels[depth + 1] = end;
    __syncthreads();
  }

  // ---- item boxes and heights, one read each, by first place
  for (int k = lane; k < count; k += BLOCK) {
    const int node = w.items[k];
    const int first = w.links[node].w >> 8;
    const TreeNode& n = gn[node];
    w.itemBox[first] = make_float4(n.box.lo.x, n.box.lo.y, n.box.hi.x, n.box.hi.y);
    w.itemHeight[first] = n.height;
  }
  __syncthreads();

  // ---- refit, deepest ranges first: boxes (cx/cy/seg/pre as float4 per range), heights
  float4* rbox = reinterpret_cast<float4*>(w.cx);
  for (int d = depth - 1; d >= 0; --d) {
    for (int r = w.levels[d] + lane; r < w.levels[d + 1]; r += BLOCK) {
      const int c[2] = {w.left[r], w.right[r]};
      AABB b[2];
      int h[2], id[2];
      bool valid = true;
      for (int e = 0; e < 2; ++e) {
        if (c[e] >= 0) {
          const float4 v = rbox[c[e]];
          b[e] = AABB{{v.x, v.y}, {v.z, v.w}};
          h[e] = w.partner[c[e]];
          id[e] = w.ranges[c[e]].w;
        } else {
          id[e] = -1 - c[e];
          if (id[e] < 0 || id[e] >= F) {
            valid = false;
            continue;
          }
          const int first = w.links[id[e]].w >> 8;
          const float4 v = w.itemBox[first];
          b[e] = AABB{{v.x, v.y}, {v.z, v.w}};
          h[e] = w.itemHeight[first];
        }
      }
      if (!valid || id[0] < 0 || id[0] >= F || id[1] < 0 || id[1] >= F) {
        bad = true;
        continue;
      }
      const AABB box = aabbUnion(b[0], b[1]);
      const int height = 1 + maxi(h[0], h[1]);
      rbox[r] = make_float4(box.lo.x, box.lo.y, box.hi.x, box.hi.y);
      w.partner[r] = height;
      const int node = w.ranges[r].w;
      TreeNode& n = gn[node];
      n.box = box;
      n.child1 = id[0];
      n.child2 = id[1];
      n.height = (uint16_t)height;
      n.flags = TN_ALLOCATED;
      gn[id[0]].parent = node;
      gn[id[1]].parent = node;
    }
    __syncthreads();
  }
  if (__syncthreads_or(bad)) {
    if (lane == 0) hd->error = 2;
    return;
  }
  if (lane == 0) {
    const int newRoot = w.ranges[0].w;
    gn[newRoot].parent = -1;
    hd->root = newRoot;
  }

  // ---- ranks: the query visits the items from their last final place to the first,
  // each item's leaves in their old order. partner: final place of each first place;
  // left: leaves at each final place; right: leaves at later final places
  for (int k = lane; k < count; k += BLOCK) {
    const int first = w.links[w.items[k]].w >> 8;
    w.partner[first] = k;
    w.left[k] = w.leaves[first];
  }
  __syncthreads();
  int carry = 0;
  for (int base = count - 1; base >= 0; base -= BLOCK) {
    const int k = base - lane;
    const int v = k >= 0 ? w.left[k] : 0;
    int total;
    const int incl = blockInclusiveAdd<BLOCK>(v, w.sums, &total);
    if (k >= 0) w.right[k] = carry + incl - v;
    carry += total;
  }
  __syncthreads();
  // owner of every old rank: the item whose interval starts at the largest start <= rank
  int run = -1;
  for (int base = 0; base < P; base += BLOCK) {
    const int r = base + lane;
    const int v = (r < P && w.lowMark[r] >= 0) ? r : -1;
    int all;
    const int incl = max(blockInclusiveMax<BLOCK>(v, w.sums, &all), run);
    if (r < P) w.byRank[r] = incl;
    run = max(run, all);
  }
  __syncthreads();
  for (int s = lane; s < shapeCount; s += BLOCK) {
    const int key = shapes[s].proxyKey;
    if (key < 0 || (key & 3) != treeType) continue;
    const int r = ranks[s];
    if (r < 0 || r >= P) continue;
    const int lo = w.byRank[r];
    if (lo < 0) continue;
    ranks[s] = w.right[w.partner[w.lowMark[lo]]] + (r - lo);
  }
  __syncthreads();
}

}  // namespace f2d
