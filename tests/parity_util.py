"""Comparison helpers for parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

from forge2d_amd import snapshot as S

BODY_FIELDS = ["px", "py", "qc", "qs", "cx", "cy", "vx", "vy", "w"]


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    den = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / den))


def body_errors(mine: S.Snapshot, ref: S.Snapshot, fields=BODY_FIELDS, floor=1.0) -> dict:
    """Max relative error per field; the denominator is max(|ref|, floor) so
    near-zero quantities are compared in absolute terms at the floor's scale."""
    return {f: rel_err(mine.bodies[f], ref.bodies[f], floor) for f in fields}


def body_mismatch_count(mine: S.Snapshot, ref: S.Snapshot, fields=BODY_FIELDS) -> int:
    return int(sum((mine.bodies[f] != ref.bodies[f]).sum() for f in fields))


def manifold_report(mine: S.Snapshot, ref: S.Snapshot, value_fields=("ax", "ay", "bx", "by", "sep"),
                    with_impulses=False) -> dict:
    """Compare manifolds keyed by (shape_a, shape_b). Returns counts of pairs
    missing on either side, point-count and feature-id mismatches, and the
    number of points whose geometry differs in any bit."""
    a = mine.contact_by_pair()
    b = ref.contact_by_pair()
    only_mine = sorted(set(a) - set(b))
    only_ref = sorted(set(b) - set(a))
    pc = ids = vals = imp = 0
    normal = 0
    for k in set(a) & set(b):
        x, y = a[k], b[k]
        n = int(y["point_count"])
        if int(x["point_count"]) != n:
            pc += 1
            continue
        if n and (x["nx"] != y["nx"] or x["ny"] != y["ny"]):
            normal += 1
        for p in range(n):
            if int(x["pts"][p]["id"]) != int(y["pts"][p]["id"]):
                ids += 1
            if any(x["pts"][p][f] != y["pts"][p][f] for f in value_fields):
                vals += 1
            if with_impulses and any(x["pts"][p][f] != y["pts"][p][f] for f in ("ni", "ti", "tni", "nv")):
                imp += 1
    return {"only_mine": only_mine, "only_ref": only_ref, "point_count": pc, "ids": ids, "values": vals,
            "normals": normal, "impulses": imp, "common": len(set(a) & set(b))}


def coloring_violations(snap: S.Snapshot) -> int:
    """Number of (colour, dynamic body) pairs that appear in more than one
    constraint of a colour (overflow excluded) plus static contacts in colour 0."""
    bodies = snap.bodies
    shapes = snap.shapes
    bad = 0
    for c in range(S.OVERFLOW):
        seen = set()
        for cid in snap.colors[c]:
            rec = snap.contacts[int(cid)]
            for sh in (int(rec["shape_a"]), int(rec["shape_b"])):
                body = int(shapes[sh]["body"])
                if int(bodies[body]["type"]) == 0:
                    if c == 0:
                        bad += 1
                    continue
                if body in seen:
                    bad += 1
                seen.add(body)
    return bad


def color_map(snap: S.Snapshot) -> dict:
    out = {}
    for c in range(S.COLOR_COUNT):
        for cid in snap.colors[c]:
            rec = snap.contacts[int(cid)]
            out[(int(rec["shape_a"]), int(rec["shape_b"]))] = c
    return out


def island_differences(mine: S.Snapshot, ref: S.Snapshot) -> list:
    """Names of the island-graph and solver-set quantities that differ between two
    version-2 snapshots: each body's set and index in it and its sleep timer; the
    island records (union-find parents, body / contact / joint lists, counts and
    removed-constraint counters), the island and solver-set id pools, the awake
    set's island and body orders, the split candidate, and the island links of
    bodies, live contacts and joints, with the joints' colours."""
    out = []
    for f in ("set_index", "local_index", "sleep_time"):
        if not np.array_equal(mine.bodies[f], ref.bodies[f]):
            out.append("body." + f)
    a, b = mine.islands, ref.islands
    if a is None or b is None:
        return out + (["island section"] if (a is None) != (b is None) else [])
    for f in ("island_count", "awake_island_count", "split_island", "set_count", "awake_body_count"):
        if getattr(a, f) != getattr(b, f):
            out.append(f)
    for f in ("islands", "free_islands", "awake_islands", "free_sets", "awake_bodies", "body_links", "joint_links"):
        if not np.array_equal(getattr(a, f), getattr(b, f)):
            out.append(f)
    live = ref.contacts["id"] >= 0
    if len(a.contact_links) != len(b.contact_links) or not np.array_equal(a.contact_links[live],
                                                                           b.contact_links[live]):
        out.append("contact_links")
    for f in ("color", "local_index"):
        if not np.array_equal(mine.joints[f], ref.joints[f]):
            out.append("joint." + f)
    return out


def tree_differences(mine: S.Snapshot, ref: S.Snapshot, box_tol: float = 0.0) -> list:
    """Names of the broad-phase tree quantities that differ between two version-3
    snapshots, per tree (static, kinematic, dynamic): root, node and proxy counts,
    the node free list (freed stack and fresh index), every node's links, height,
    flags and category bits, and its box (exactly, or within ``box_tol`` absolute)."""
    out = []
    if mine.trees is None or ref.trees is None:
        return ["tree section"] if (mine.trees is None) != (ref.trees is None) else []
    for k, (a, b) in enumerate(zip(mine.trees, ref.trees)):
        name = ("static", "kinematic", "dynamic")[k]
        for f in ("root", "node_count", "proxy_count", "fresh"):
            if getattr(a, f) != getattr(b, f):
                out.append(f"{name}.{f}")
        if not np.array_equal(a.freed, b.freed):
            out.append(f"{name}.freed")
        if len(a.nodes) != len(b.nodes):
            out.append(f"{name}.nodes")
            continue
        for f in ("child1", "child2", "parent", "height", "flags", "cat_lo", "cat_hi"):
            if not np.array_equal(a.nodes[f], b.nodes[f]):
                out.append(f"{name}.{f}")
        d = np.abs(a.nodes["aabb"].astype(np.float64) - b.nodes["aabb"].astype(np.float64))
        if d.max(initial=0.0) > box_tol:
            out.append(f"{name}.aabb")
    return out


QUERY_FIXTURES = ["queries", "queries_sleep", "queries_wake2"]


def query_mismatches(queries, expected, counts, hits) -> list:
    """Indices of the queries whose results differ from the expected ones (a list of
    hit arrays per query), in order: every hit's shape, point, normal and fraction
    bit for bit, in the order the reference's tree walks report them."""
    bad = []
    for i in range(len(queries)):
        ref = np.asarray(expected[i])
        mine = np.asarray(hits[i][:counts[i]]) if counts[i] > 0 else hits[i][:0]
        if counts[i] != len(ref):
            bad.append(i)
            continue
        for f in ref.dtype.names:
            if not np.array_equal(ref[f].view(np.uint32), mine[f].view(np.uint32)):
                bad.append(i)
                break
    return bad


def _h(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


DIGEST_BODY_FIELDS = ["type", "set_index", "flags", "px", "py", "qc", "qs", "cx", "cy", "c0x", "c0y", "q0c", "q0s",
                      "vx", "vy", "w", "inv_mass", "inv_i", "min_ext", "max_ext", "sleep_time"]
DIGEST_CONTACT_FIELDS = ["id", "shape_a", "shape_b", "flags", "color", "sim_flags", "point_count", "friction",
                         "restitution", "rolling", "tangent_speed", "nx", "ny", "rolling_impulse", "pts"]


def state_digest(s: S.Snapshot) -> dict:
    """Hashes of one snapshot's state, field by field, for comparing a large world
    with a reference run without storing the reference's snapshots: bodies, shape
    boxes and proxy keys, live contacts by id with their manifolds and impulses,
    colour arrays, the free-id list, the move array and the trees."""
    d = {"step": int(s.step), "bodies": int(len(s.bodies)), "live_contacts": int(len(s.live_contacts())),
         "fast_bodies": int(((s.bodies["flags"] & 1) != 0).sum())}
    for f in DIGEST_BODY_FIELDS:
        d["body." + f] = _h(s.bodies[f])
    for f in ("aabb", "fat", "proxy_key"):
        d["shape." + f] = _h(s.shapes[f])
    live = np.sort(s.live_contacts(), order="id")
    for f in DIGEST_CONTACT_FIELDS:
        d["contact." + f] = _h(live[f])
    d["colors"] = _h(np.concatenate([np.asarray(c, np.int32) for c in s.colors] +
                                    [np.array([len(c) for c in s.colors], np.int32)]))
    d["free_ids"] = _h(np.asarray(s.free_ids, np.int32))
    d["move"] = _h(np.asarray(s.move, np.int32))
    if s.trees is not None:
        for k, t in enumerate(s.trees):
            name = ("static", "kinematic", "dynamic")[k]
            d[f"tree.{name}"] = _h(np.array([t.root, t.node_count, t.proxy_count, t.fresh], np.int32))
            d[f"tree.{name}.freed"] = _h(np.asarray(t.freed, np.int32))
            for f in ("child1", "child2", "parent", "height", "flags", "aabb"):
                d[f"tree.{name}.{f}"] = _h(t.nodes[f])
    return d


def digest_differences(mine: dict, ref: dict) -> list:
    """Names of the digest entries that differ."""
    return [k for k in ref if mine.get(k) != ref[k]]
