"""World-state snapshot format (checkpoint / parity fixtures).

One snapshot holds the complete persistent state of one world after a step:
bodies, shapes (AABB, fat AABB, proxy key), contacts by contact id (manifold,
colour, flags), revolute joints, the constraint-graph colour arrays in array
order, the broad-phase move array and the contact-id free list. The engine's
``f2dg_export_world`` writes it and ``f2dg_import_world`` reads it; the
reference fixture harness (``oracle/ref_harness.c``) writes the same layout
from the reference's internal structures.

Records (little-endian, packed):
  HEADER 64 B, BODY 112 B, SHAPE 48 B, CONTACT 168 B, JOINT 64 B, then
  int32 colour counts[12], int32 colour ids[sum], int32 move keys[n],
  int32 free contact ids[n].
Version 2 appends the island section (island graph and solver sets):
  int32 header[16] = [island array count, free island ids, awake islands,
  split island id, solver-set array count, free set ids, awake bodies, island
  id-pool next index, set id-pool next index, 0...]; ISLAND 64 B records; int32
  free island ids; int32 awake island ids (awake set order); int32 free set ids;
  int32 awake body ids (awake set order); per body int32[4] (island id, prev,
  next, head contact key); per contact slot int32[8] (island id, prev, next, set
  index, body-list edge keys prevA, nextA, prevB, nextB); per joint int32[4]
  (island id, prev, next, set index).
Version 3 appends the broad-phase trees (static, kinematic, dynamic; the
reference's b2DynamicTree, S/dynamic_tree.c): per tree int32 header[8] = [root,
node count, proxy count, fresh index, freed count, 0, 0, 0], TREE_NODE 40 B records
for node ids [0, fresh index) (free nodes zeroed with -1 links), then the freed node
ids as a stack, bottom first (the next allocation pops the last; once empty, fresh
ids are taken in ascending order).
"""
from __future__ import annotations

import numpy as np

SNAP_MAGIC = 0x50414E53
COLOR_COUNT = 12
OVERFLOW = 11

HEADER_DT = np.dtype([
    ("magic", "<i4"), ("version", "<i4"), ("step", "<i4"), ("body_count", "<i4"), ("shape_count", "<i4"),
    ("contact_cap", "<i4"), ("joint_count", "<i4"), ("overflow_count", "<i4"), ("move_count", "<i4"),
    ("free_id_count", "<i4"), ("color_total", "<i4"), ("reserved0", "<i4"),
    ("contact_speed", "<f4"), ("inv_h", "<f4"), ("reserved1", "<f4"), ("reserved2", "<f4"),
])
BODY_DT = np.dtype([
    ("type", "<i4"), ("set_index", "<i4"), ("local_index", "<i4"), ("flags", "<i4"),
    ("px", "<f4"), ("py", "<f4"), ("qc", "<f4"), ("qs", "<f4"), ("cx", "<f4"), ("cy", "<f4"),
    ("c0x", "<f4"), ("c0y", "<f4"), ("q0c", "<f4"), ("q0s", "<f4"), ("lcx", "<f4"), ("lcy", "<f4"),
    ("vx", "<f4"), ("vy", "<f4"), ("w", "<f4"), ("fx", "<f4"), ("fy", "<f4"), ("torque", "<f4"),
    ("inv_mass", "<f4"), ("inv_i", "<f4"), ("min_ext", "<f4"), ("max_ext", "<f4"), ("sleep_time", "<f4"),
    ("reserved", "<f4"),
])
SHAPE_DT = np.dtype([
    ("aabb", "<f4", (4,)), ("fat", "<f4", (4,)), ("enlarged", "<i4"), ("proxy_key", "<i4"), ("body", "<i4"),
    ("reserved", "<i4"),
])
POINT_DT = np.dtype([
    ("px", "<f4"), ("py", "<f4"), ("ax", "<f4"), ("ay", "<f4"), ("bx", "<f4"), ("by", "<f4"),
    ("sep", "<f4"), ("ni", "<f4"), ("ti", "<f4"), ("tni", "<f4"), ("nv", "<f4"), ("id", "<i4"),
    ("persisted", "<i4"),
])
CONTACT_DT = np.dtype([
    ("id", "<i4"), ("shape_a", "<i4"), ("shape_b", "<i4"), ("flags", "<i4"), ("color", "<i4"),
    ("local_index", "<i4"), ("sim_flags", "<i4"), ("point_count", "<i4"), ("friction", "<f4"),
    ("restitution", "<f4"), ("rolling", "<f4"), ("tangent_speed", "<f4"), ("nx", "<f4"), ("ny", "<f4"),
    ("rolling_impulse", "<f4"), ("reserved", "<f4"), ("pts", POINT_DT, (2,)),
])
JOINT_DT = np.dtype([
    ("id", "<i4"), ("color", "<i4"), ("local_index", "<i4"), ("body_a", "<i4"), ("body_b", "<i4"),
    ("enable_limit", "<i4"), ("enable_motor", "<i4"), ("enable_spring", "<i4"),
    ("linear_impulse_x", "<f4"), ("linear_impulse_y", "<f4"), ("spring_impulse", "<f4"),
    ("motor_impulse", "<f4"), ("lower_impulse", "<f4"), ("upper_impulse", "<f4"), ("ref_angle", "<f4"),
    ("reserved", "<f4"),
])
ISLAND_DT = np.dtype([
    ("id", "<i4"), ("set_index", "<i4"), ("local_index", "<i4"), ("head_body", "<i4"), ("tail_body", "<i4"),
    ("body_count", "<i4"), ("head_contact", "<i4"), ("tail_contact", "<i4"), ("contact_count", "<i4"),
    ("head_joint", "<i4"), ("tail_joint", "<i4"), ("joint_count", "<i4"), ("parent", "<i4"),
    ("remove_count", "<i4"), ("reserved", "<i4", (2,)),
])
LINK_DT = np.dtype([("island", "<i4"), ("prev", "<i4"), ("next", "<i4"), ("set_index", "<i4")])
BODY_LINK_DT = np.dtype([("island", "<i4"), ("prev", "<i4"), ("next", "<i4"), ("head_contact", "<i4")])
CONTACT_LINK_DT = np.dtype([("island", "<i4"), ("prev", "<i4"), ("next", "<i4"), ("set_index", "<i4"),
                            ("prev_a", "<i4"), ("next_a", "<i4"), ("prev_b", "<i4"), ("next_b", "<i4")])
# b2TreeNode, S/dynamic_tree.c:19-50: box, category bits, children (leaf: user data =
# shape id in child1), parent, height, flags (allocated 1, enlarged 2, leaf 4)
TREE_NODE_DT = np.dtype([
    ("aabb", "<f4", (4,)), ("cat_lo", "<u4"), ("cat_hi", "<u4"), ("child1", "<i4"), ("child2", "<i4"),
    ("parent", "<i4"), ("height", "<u2"), ("flags", "<u2"),
])
TREE_ALLOCATED, TREE_ENLARGED, TREE_LEAF = 1, 2, 4
assert TREE_NODE_DT.itemsize == 40
assert ISLAND_DT.itemsize == 64 and LINK_DT.itemsize == 16 and CONTACT_LINK_DT.itemsize == 32
assert HEADER_DT.itemsize == 64 and BODY_DT.itemsize == 112 and SHAPE_DT.itemsize == 48
assert POINT_DT.itemsize == 52 and CONTACT_DT.itemsize == 168 and JOINT_DT.itemsize == 64

# b2ContactFlags / b2ContactSimFlags (S/contact.h)
CONTACT_TOUCHING = 0x1
SIM_TOUCHING = 0x00010000


class Islands:
    """Island graph and solver-set bookkeeping of a version-2 snapshot."""
    __slots__ = ("island_count", "awake_island_count", "split_island", "set_count", "awake_body_count",
                 "island_next", "set_next", "islands", "free_islands", "awake_islands", "free_sets", "awake_bodies",
                 "body_links", "contact_links", "joint_links")


class Tree:
    """One broad-phase tree of a version-3 snapshot."""
    __slots__ = ("root", "node_count", "proxy_count", "fresh", "nodes", "freed")

    def leaf_order(self) -> np.ndarray:
        """Shape ids of the leaves in the order a query visits them (child2 before
        child1, S/dynamic_tree.c:1114-1170)."""
        out = []
        if self.root < 0:
            return np.zeros(0, np.int32)
        stack = [self.root]
        while stack:
            n = stack.pop()
            rec = self.nodes[n]
            if rec["flags"] & TREE_LEAF:
                out.append(int(rec["child1"]))
            else:
                stack.append(int(rec["child1"]))
                stack.append(int(rec["child2"]))
        return np.array(out, np.int32)


class Snapshot:
    __slots__ = ("header", "bodies", "shapes", "contacts", "joints", "colors", "move", "free_ids", "islands",
                 "trees", "nbytes")

    @property
    def step(self) -> int:
        return int(self.header["step"])

    def live_contacts(self) -> np.ndarray:
        return self.contacts[self.contacts["id"] >= 0]

    def pair_keys(self) -> np.ndarray:
        """(shape_a, shape_b) of every live contact as int64 keys a<<32|b, sorted."""
        c = self.live_contacts()
        return np.sort((c["shape_a"].astype(np.int64) << 32) | c["shape_b"].astype(np.int64))

    def contact_by_pair(self) -> dict:
        c = self.live_contacts()
        return {(int(a), int(b)): rec for a, b, rec in zip(c["shape_a"], c["shape_b"], c)}


def parse_one(buf, off: int = 0) -> Snapshot:
    s = Snapshot()
    h = np.frombuffer(buf, HEADER_DT, 1, off)[0]
    if int(h["magic"]) != SNAP_MAGIC:
        raise ValueError("bad snapshot magic at offset %d" % off)
    o = off + HEADER_DT.itemsize
    s.header = h
    s.bodies = np.frombuffer(buf, BODY_DT, int(h["body_count"]), o)
    o += s.bodies.nbytes
    s.shapes = np.frombuffer(buf, SHAPE_DT, int(h["shape_count"]), o)
    o += s.shapes.nbytes
    s.contacts = np.frombuffer(buf, CONTACT_DT, int(h["contact_cap"]), o)
    o += s.contacts.nbytes
    s.joints = np.frombuffer(buf, JOINT_DT, int(h["joint_count"]), o)
    o += s.joints.nbytes
    counts = np.frombuffer(buf, "<i4", COLOR_COUNT, o)
    o += 4 * COLOR_COUNT
    ids = np.frombuffer(buf, "<i4", int(counts.sum()), o)
    o += ids.nbytes
    s.colors = []
    k = 0
    for c in range(COLOR_COUNT):
        s.colors.append(ids[k:k + int(counts[c])])
        k += int(counts[c])
    s.move = np.frombuffer(buf, "<i4", int(h["move_count"]), o)
    o += s.move.nbytes
    s.free_ids = np.frombuffer(buf, "<i4", int(h["free_id_count"]), o)
    o += s.free_ids.nbytes
    s.islands = None
    s.trees = None
    if int(h["version"]) >= 2:
        s.islands, o = _parse_islands(buf, o, len(s.bodies), len(s.contacts), len(s.joints))
    if int(h["version"]) >= 3:
        s.trees = []
        for _ in range(3):
            t, o = _parse_tree(buf, o)
            s.trees.append(t)
    s.nbytes = o - off
    return s


def _parse_tree(buf, o):
    hd = np.frombuffer(buf, "<i4", 8, o)
    o += 32
    t = Tree()
    t.root, t.node_count, t.proxy_count, t.fresh = (int(x) for x in hd[:4])
    t.nodes = np.frombuffer(buf, TREE_NODE_DT, t.fresh, o)
    o += t.nodes.nbytes
    t.freed = np.frombuffer(buf, "<i4", int(hd[4]), o)
    o += t.freed.nbytes
    return t, o


def _tree_bytes(t: Tree) -> bytes:
    hd = np.zeros(8, "<i4")
    hd[:5] = [t.root, t.node_count, t.proxy_count, t.fresh, len(t.freed)]
    return hd.tobytes() + np.ascontiguousarray(t.nodes).tobytes() + np.asarray(t.freed, "<i4").tobytes()


def _parse_islands(buf, o, nb, nc, nj):
    hd = np.frombuffer(buf, "<i4", 16, o)
    o += 64
    isl = Islands()
    isl.island_count, nfree, isl.awake_island_count, isl.split_island = (int(x) for x in hd[:4])
    isl.set_count, nsetfree, isl.awake_body_count, isl.island_next, isl.set_next = (int(x) for x in hd[4:9])
    isl.islands = np.frombuffer(buf, ISLAND_DT, isl.island_count, o)
    o += isl.islands.nbytes
    for name, n in (("free_islands", nfree), ("awake_islands", isl.awake_island_count), ("free_sets", nsetfree),
                    ("awake_bodies", isl.awake_body_count)):
        a = np.frombuffer(buf, "<i4", n, o)
        o += a.nbytes
        setattr(isl, name, a)
    for name, n, dt in (("body_links", nb, BODY_LINK_DT), ("contact_links", nc, CONTACT_LINK_DT),
                        ("joint_links", nj, LINK_DT)):
        a = np.frombuffer(buf, dt, n, o)
        o += a.nbytes
        setattr(isl, name, a)
    return isl, o


def parse_all(buf) -> list[Snapshot]:
    out = []
    off = 0
    while off < len(buf):
        s = parse_one(buf, off)
        out.append(s)
        off += s.nbytes
    return out


def load(path: str) -> list[Snapshot]:
    with open(path, "rb") as f:
        return parse_all(f.read())


def created_pairs(before: Snapshot, after: Snapshot) -> list[tuple[int, int]]:
    """Contacts alive in ``after`` but not in ``before`` (matched by id and shape
    pair), in creation order. Creation pops the LIFO contact-id pool
    (S/id_pool.c:19-37): first the free list from its top, then fresh ids."""
    old = {}
    for c in before.live_contacts():
        old[int(c["id"])] = (int(c["shape_a"]), int(c["shape_b"]))
    new = {}
    for c in after.live_contacts():
        key = (int(c["shape_a"]), int(c["shape_b"]))
        if old.get(int(c["id"])) != key:
            new[int(c["id"])] = key
    order = [int(i) for i in before.free_ids[::-1]]
    nxt = len(before.contacts)
    out = []
    pool = list(order)
    # ids are handed out in pop order; contacts destroyed in the same step free
    # ids only after creation, so the pop sequence is exactly pool then fresh.
    seq = pool + list(range(nxt, nxt + len(new) + 1))
    for cid in seq:
        if cid in new:
            out.append(new[cid])
            if len(out) == len(new):
                break
    return out


def to_bytes(s: Snapshot) -> bytes:
    """Inverse of ``parse_one``."""
    h = np.array(s.header, HEADER_DT).copy()
    h["body_count"] = len(s.bodies)
    h["shape_count"] = len(s.shapes)
    h["contact_cap"] = len(s.contacts)
    h["joint_count"] = len(s.joints)
    h["overflow_count"] = len(s.colors[OVERFLOW])
    h["move_count"] = len(s.move)
    h["free_id_count"] = len(s.free_ids)
    h["color_total"] = int(sum(len(c) for c in s.colors))
    parts = [h.tobytes(), np.ascontiguousarray(s.bodies).tobytes(), np.ascontiguousarray(s.shapes).tobytes(),
             np.ascontiguousarray(s.contacts).tobytes(), np.ascontiguousarray(s.joints).tobytes(),
             np.array([len(c) for c in s.colors], "<i4").tobytes()]
    parts += [np.asarray(c, "<i4").tobytes() for c in s.colors]
    parts += [np.asarray(s.move, "<i4").tobytes(), np.asarray(s.free_ids, "<i4").tobytes()]
    isl = getattr(s, "islands", None)
    trees = getattr(s, "trees", None)
    h["version"] = (3 if trees is not None else 2) if isl is not None else 1
    parts[0] = h.tobytes()
    if isl is not None:
        hd = np.zeros(16, "<i4")
        hd[:9] = [isl.island_count, len(isl.free_islands), isl.awake_island_count, isl.split_island, isl.set_count,
                  len(isl.free_sets), isl.awake_body_count, isl.island_next, isl.set_next]
        parts.append(hd.tobytes())
        for a in (isl.islands, isl.free_islands, isl.awake_islands, isl.free_sets, isl.awake_bodies,
                  isl.body_links, isl.contact_links, isl.joint_links):
            parts.append(np.ascontiguousarray(a).tobytes())
        if trees is not None:
            parts += [_tree_bytes(t) for t in trees]
    return b"".join(parts)


def copy(s: Snapshot) -> Snapshot:
    o = Snapshot()
    o.header = np.array(s.header, HEADER_DT).copy()
    o.bodies = s.bodies.copy()
    o.shapes = s.shapes.copy()
    o.contacts = s.contacts.copy()
    o.joints = s.joints.copy()
    o.colors = [np.array(c, "<i4") for c in s.colors]
    o.move = np.array(s.move, "<i4")
    o.free_ids = np.array(s.free_ids, "<i4")
    o.islands = getattr(s, "islands", None)
    o.trees = getattr(s, "trees", None)
    o.nbytes = s.nbytes
    return o


def warm_start_matched(prev: Snapshot, cur: Snapshot) -> np.ndarray:
    """Contacts of ``cur`` with the impulses they carried into the solver of the
    step that produced ``cur``: manifold geometry of ``cur``, warm-start impulses
    matched by feature id from the manifold of ``prev`` (b2UpdateContact,
    S/contact.c:555-600)."""
    out = cur.contacts.copy()
    old = prev.contact_by_pair()
    for i in range(len(out)):
        rec = out[i]
        if rec["id"] < 0:
            continue
        key = (int(rec["shape_a"]), int(rec["shape_b"]))
        n = int(rec["point_count"])
        o = old.get(key)
        if o is not None:
            o = o.copy()
        if n > 0:
            rec["rolling_impulse"] = o["rolling_impulse"] if o is not None else 0.0
        for p in range(n):
            pt = rec["pts"][p]
            pt["ni"] = 0.0
            pt["ti"] = 0.0
            pt["tni"] = 0.0
            pt["nv"] = 0.0
            pt["persisted"] = 0
            if o is None:
                continue
            for q in range(int(o["point_count"])):
                if int(o["pts"][q]["id"]) == int(pt["id"]):
                    pt["ni"] = o["pts"][q]["ni"]
                    pt["ti"] = o["pts"][q]["ti"]
                    pt["persisted"] = 1
                    o["pts"][q]["ni"] = 0.0
                    o["pts"][q]["ti"] = 0.0
                    break
        out[i] = rec
    return out


def post_collide(prev: Snapshot, cur: Snapshot) -> Snapshot:
    """The world as the solver of step ``cur.step`` saw it: bodies and shapes from
    ``prev`` (finalize of the previous step), contacts, colours and the overflow
    order from ``cur`` with warm-start impulses matched from ``prev``."""
    s = copy(cur)
    s.bodies = prev.bodies.copy()
    s.shapes = prev.shapes.copy()
    s.contacts = warm_start_matched(prev, cur)
    s.joints = prev.joints.copy()
    s.joints["color"] = cur.joints["color"]
    s.move = np.zeros(0, "<i4")
    s.header["step"] = prev.header["step"]
    # the island graph after the collide is not part of either snapshot: the
    # composed state carries none (version 1), for worlds with sleep off
    s.islands = None
    s.trees = None
    return s
