"""World queries on the device (queryKernel and explodeKernel behind f2dg_query, and
the flat f2d_world_cast_ray_closest / f2d_world_cast_ray / f2d_world_overlap_aabb /
f2d_world_explode) against the reference and the oracle:

* the reference's own results on the query fixtures (tests/golden, written by the
  reference build: b2World_CastRayClosest, b2World_CastRay, b2World_OverlapAABB,
  b2World_Explode, S/world.c:2050-2308, 2640-2747) from the reference's snapshot:
  every hit bit for bit and in the reference's tree-walk order, and after the
  explosions every body, the island graph, the solver sets (in the order the
  explosions woke them) and the colouring;
* the oracle (oracle/b2oracle.c, pinned by tests/test_oracle_queries.py) on larger
  scenes the device has stepped, and on a batch of 64 mixed worlds queried in one
  launch (worlds interleaved, two explosions in one world run in order);
* the reference's own query tests, F/test/api/events_test.dart:151-265, replayed
  through the flat ABI with their tolerances.
"""
import ctypes
import os

import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes
from forge2d_amd import snapshot as S
from oracle import pyoracle as O
from tests import parity_util as P
from tests.test_gpu_api import body, box, shape_def, vel, world
from tests.test_gpu_oracle import HARNESS, SCENES, reference_geometry
from tests.test_oracle_queries import check_after_explosions, query_fixture

pytestmark = pytest.mark.gpu
needs_oracle = pytest.mark.skipif(not (O.available("f32") and os.path.exists(HARNESS)),
                                  reason="oracle or reference build missing")
ALL = F.ALL_BITS


@pytest.mark.parametrize("name", P.QUERY_FIXTURES)
def test_device_queries_match_reference(gpu, name):
    spec, geoms, snaps, raws, q, ref = query_fixture(name)
    print(f"[queries] fixture {name}: {len(q)} queries", flush=True)
    b = F.Batch(spec)
    b.import_state(0, raws[0])
    counts, hits = b.query(q, cap=256)
    assert P.query_mismatches(q, ref, counts, hits) == []
    check_after_explosions(name, S.parse_one(b.export(0)), snaps[1])
    b.close()


def random_queries(snap, seed, n_rays, n_boxes, explosions, world=0, extent=None):
    """Seeded rays (a quarter from inside bodies), boxes and explosions over the scene."""
    rng = np.random.default_rng(seed)
    px, py = snap.bodies["px"], snap.bodies["py"]
    lo = np.array([px.min() - 2.0, py.min() - 2.0]) if extent is None else np.array(extent[0])
    hi = np.array([px.max() + 2.0, py.max() + 2.0]) if extent is None else np.array(extent[1])
    size = float(np.max(hi - lo))
    rows = []
    for i in range(n_rays):
        if i % 4 == 0:
            k = rng.integers(len(px))
            o = (float(px[k]), float(py[k]))
        else:
            o = tuple(float(v) for v in rng.uniform(lo, hi))
        a = rng.uniform(-np.pi, np.pi)
        ln = rng.uniform(0.1, 1.0) * size
        kind = F.Q_RAY_CLOSEST if i % 2 == 0 else F.Q_RAY_ALL
        rows.append((world, kind, [o[0], o[1], ln * np.cos(a), ln * np.sin(a)]))
    for _ in range(n_boxes):
        c = rng.uniform(lo, hi)
        e = rng.uniform(0.05, 0.2, 2) * size
        rows.append((world, F.Q_OVERLAP_AABB, [c[0] - e[0], c[1] - e[1], c[0] + e[0], c[1] + e[1]]))
    for _ in range(explosions):
        k = rng.integers(len(px))
        rows.append((world, F.Q_EXPLODE, [float(px[k]), float(py[k]), rng.uniform(0.5, 3.0), rng.uniform(0.0, 2.0),
                                          rng.uniform(0.2, 2.0)], (0, 0, ALL, ALL)))
    return F.make_queries(rows)


QUERY_SCENES = ["pyramid48", "container400_ccd", "container2600"]


@needs_oracle
@pytest.mark.parametrize("name", QUERY_SCENES)
def test_device_queries_match_oracle_on_stepped_scenes(gpu, name):
    spec = scenes.to_bytes([SCENES[name]()])
    geoms = reference_geometry(spec)
    b = F.Batch(spec)
    b.step(steps=30)
    assert b.sync() == 0
    raw = b.export(0)
    snap = S.parse_one(raw)
    q = random_queries(snap, 11, 1500, 300, 4)
    print(f"[queries] {name}: {len(q)} queries after 30 steps", flush=True)
    cap = 1024
    counts, hits = b.query(q, cap=cap)
    o = O.Oracle(spec, geoms, raw)
    oc, oh = o.query(q, cap=cap)
    assert (counts >= 0).all() and counts.max() < cap
    assert sum(int(c) for c in counts[q["kind"] != F.Q_EXPLODE]) > 500
    exp = [oh[i][:oc[i]] for i in range(len(q))]
    assert P.query_mismatches(q, exp, counts, hits) == []
    mine, ref = S.parse_one(b.export(0)), S.parse_one(o.export())
    for f in P.BODY_FIELDS:
        assert np.array_equal(mine.bodies[f], ref.bodies[f]), f
    assert (mine.bodies["vx"] != snap.bodies["vx"]).any()
    b.close()


@needs_oracle
def test_queries_over_a_batch_of_worlds(gpu):
    worlds = scenes.config4(64, lo=10, hi=200)
    spec = scenes.to_bytes(worlds)
    b = F.Batch(spec)
    b.step(steps=40)
    assert b.sync() == 0
    raws = [b.export(i) for i in range(len(worlds))]
    per_world = []
    for i in range(len(worlds)):
        per_world.append(random_queries(S.parse_one(raws[i]), 100 + i, 24, 6, 2 if i % 7 == 0 else 1, world=i))
    # interleave the worlds' queries; two explosions of one world run in their order
    keys = np.concatenate([np.arange(len(p)) * len(worlds) + i for i, p in enumerate(per_world)])
    q = np.concatenate(per_world)[np.argsort(keys, kind="stable")]
    bad = F.make_queries([(len(worlds), F.Q_RAY_CLOSEST, [0.0, 0.0, 1.0, 0.0]), (-1, F.Q_OVERLAP_AABB, [0, 0, 1, 1])])
    q = np.concatenate([q, bad])
    print(f"[queries] batch of {len(worlds)} worlds: {len(q)} queries in one call", flush=True)
    counts, hits = b.query(q, cap=256)
    assert list(counts[-2:]) == [-1, -1]
    for i in range(len(worlds)):
        sel = np.nonzero(q["world"] == i)[0]
        geoms = reference_geometry_world(spec, i)
        o = O.Oracle(spec, geoms, raws[i], world=i)
        oc, oh = o.query(q[sel], cap=256)
        exp = [oh[k][:oc[k]] for k in range(len(sel))]
        assert P.query_mismatches(q[sel], exp, counts[sel], hits[sel]) == [], i
        mine, ref = S.parse_one(b.export(i)), S.parse_one(o.export())
        for f in P.BODY_FIELDS:
            assert np.array_equal(mine.bodies[f], ref.bodies[f]), (i, f)
    b.close()


def reference_geometry_world(spec: bytes, world: int) -> bytes:
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as td:
        sp, gp = os.path.join(td, "s.bin"), os.path.join(td, "g.bin")
        with open(sp, "wb") as f:
            f.write(spec)
        subprocess.run([HARNESS, "geom", sp, gp, str(world)], check=True, timeout=60)
        with open(gp, "rb") as f:
            return f.read()


def test_query_device_entry_point_matches_host_entry_point(gpu):
    """f2dg_query_device (device arrays, one launch on the batch's stream) returns what
    f2dg_query returns."""
    import torch

    spec = scenes.to_bytes([SCENES["container400_ccd"]()])
    b = F.Batch(spec)
    b.step(steps=20)
    assert b.sync() == 0
    q = random_queries(S.parse_one(b.export(0)), 5, 400, 100, 0)
    counts, hits = b.query(q, cap=128)
    dq = torch.from_numpy(q.view(np.uint8).copy()).cuda()
    dh = torch.zeros(len(q) * 128 * 24, dtype=torch.uint8, device="cuda")
    dc = torch.zeros(len(q), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert F.lib().f2dg_query_device(b.h, dq.data_ptr(), len(q), dh.data_ptr(), dc.data_ptr(), 128) == 0
    b.synchronize()
    c2 = dc.cpu().numpy()
    h2 = dh.cpu().numpy().view(F.HIT_DTYPE).reshape(len(q), 128)
    assert np.array_equal(counts, c2)
    for i in range(len(q)):
        n = min(int(counts[i]), 128)
        assert h2[i][:n].tobytes() == hits[i][:n].tobytes(), i
    b.close()


# ---- F/test/api/events_test.dart:151-265 through the flat ABI

def circle(bid, r, category=1):
    out = (ctypes.c_int32 * 2)()
    d = shape_def()
    d.category_lo = category
    F.lib().f2d_create_circle_shape(bid[0], bid[1], ctypes.byref(d), 0.0, 0.0, r, out)
    assert out[0] > 0
    return out[0], ctypes.c_uint32(out[1]).value


def closest(w, origin, translation, category=1, mask=ALL):
    out = (ctypes.c_double * 7)()
    hit = F.lib().f2d_world_cast_ray_closest(w, origin[0], origin[1], translation[0], translation[1], category, 0,
                                             mask & 0xFFFFFFFF, ALL if mask == ALL else 0, out)
    return list(out) if hit else None


def cast_ray(w, origin, translation, fn):
    cb = F.CAST_RAY_FN(fn)
    F.lib().f2d_set_query_callbacks(ctypes.cast(cb, ctypes.c_void_p), None)
    try:
        F.lib().f2d_world_cast_ray(w, origin[0], origin[1], translation[0], translation[1], 1, 0, ALL, ALL)
    finally:
        F.lib().f2d_set_query_callbacks(None, None)


def test_cast_ray_closest_finds_the_nearest_shape(gpu):
    L = F.lib()
    w = world()
    ground = body(w, 0, 0.0, -1.0)
    box(ground, 50.0, 1.0)
    near = body(w, 0, 2.0, 0.5)
    box(near, 0.5, 0.5)
    far = body(w, 0, 4.0, 0.5)
    box(far, 0.5, 0.5)
    hit = closest(w, (0.0, 0.5), (10.0, 0.0))
    assert hit is not None
    assert int(hit[0]) == 2 and abs(hit[2] - 1.5) < 0.01 and abs(hit[4] + 1.0) < 0.01 and abs(hit[6] - 0.15) < 0.01
    assert closest(w, (0.0, 10.0), (1.0, 0.0)) is None
    L.f2d_destroy_world(w)


def test_cast_ray_all_sorted_and_callback_stops_early(gpu):
    L = F.lib()
    w = world()
    first = body(w, 0, 2.0, 0.0)
    circle(first, 0.5)
    second = body(w, 0, 5.0, 0.0)
    circle(second, 0.5)
    hits = []

    def collect(i1, wg, px, py, nx, ny, fraction):
        hits.append((i1, fraction))
        return 1.0

    cast_ray(w, (0.0, 0.0), (10.0, 0.0), collect)
    hits.sort(key=lambda h: h[1])
    assert [h[0] for h in hits] == [1, 2] and hits[0][1] < hits[1][1]
    calls = []

    def stop(i1, wg, px, py, nx, ny, fraction):
        calls.append(i1)
        return 0.0

    cast_ray(w, (0.0, 0.0), (10.0, 0.0), stop)
    assert len(calls) == 1
    L.f2d_destroy_world(w)


def test_query_filters_exclude_non_matching_categories(gpu):
    L = F.lib()
    w = world()
    b = body(w, 0, 2.0, 0.0)
    circle(b, 0.5, category=2)
    assert closest(w, (0.0, 0.0), (10.0, 0.0), mask=4) is None
    assert closest(w, (0.0, 0.0), (10.0, 0.0), mask=2) is not None
    L.f2d_destroy_world(w)


def test_overlap_aabb_returns_shapes_in_the_box(gpu):
    L = F.lib()
    w = world()
    inside = body(w, 0, 0.0, 0.0)
    sid = circle(inside, 0.5)
    outside = body(w, 0, 20.0, 0.0)
    circle(outside, 0.5)
    found = []
    cb = F.OVERLAP_FN(lambda i1, wg: found.append((i1, wg)) or 1)
    L.f2d_set_query_callbacks(None, ctypes.cast(cb, ctypes.c_void_p))
    try:
        L.f2d_world_overlap_aabb(w, -2.0, -2.0, 2.0, 2.0, 1, 0, ALL, ALL)
    finally:
        L.f2d_set_query_callbacks(None, None)
    assert found == [sid]
    L.f2d_destroy_world(w)


def test_explosion_pushes_nearby_bodies_away(gpu):
    L = F.lib()
    w = world(gravity=(0.0, 0.0))
    b = body(w, 2, 1.0, 0.0)
    circle(b, 0.5)
    L.f2d_world_explode(w, ALL, ALL, 0.0, 0.0, 3.0, 0.0, 5.0)
    L.f2d_world_step(w, 1.0 / 60.0, 4)
    assert vel(b)[0] > 0.1
    L.f2d_destroy_world(w)
