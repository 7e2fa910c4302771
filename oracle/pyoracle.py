"""ctypes front end of the C restatement in oracle/b2oracle.c (test infrastructure).

Only the tests, the smoke check and the CPU-baseline leg of bench.py use this
module, as the checker. It loads oracle/_build/liboracle_{f32,f64}.so, built by
``make -C oracle port``.

    o = Oracle(spec, geoms, snapshot_bytes, precision="f32")
    o.update_pairs()                # the pair update over the snapshot's trees (version 3)
    o.create_contacts(created)      # or: contacts the broad phase creates this step, given
    o.collide()                     # manifolds + contact state loop + colouring
    o.solve(dt, substeps)           # solver stages, finalize, continuous, moves
    o.export() -> snapshot bytes;  o.bodies() -> (n, 7) float64
    o.query(queries) -> (counts, hits)   # ray casts, AABB overlaps, explosions
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs: dict = {}


def lib_path(precision: str) -> str:
    return os.path.join(HERE, "_build", f"liboracle_{precision}.so")


def available(precision: str = "f32") -> bool:
    return os.path.exists(lib_path(precision))


def _lib(precision: str):
    if precision not in _libs:
        L = ctypes.CDLL(lib_path(precision))
        vp, L_ = ctypes.c_void_p, ctypes.c_long
        L.b2o_create.restype = vp
        L.b2o_create.argtypes = [vp, L_, ctypes.c_int, vp, L_, vp, L_]
        L.b2o_destroy.argtypes = [vp]
        L.b2o_update_pairs.restype = ctypes.c_int
        L.b2o_update_pairs.argtypes = [vp]
        L.b2o_create_contacts.restype = ctypes.c_int
        L.b2o_create_contacts.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        L.b2o_collide.restype = ctypes.c_int
        L.b2o_collide.argtypes = [vp]
        L.b2o_solve.restype = ctypes.c_int
        L.b2o_solve.argtypes = [vp, ctypes.c_float, ctypes.c_int]
        L.b2o_export.restype = L_
        L.b2o_export.argtypes = [vp, vp, L_]
        L.b2o_bodies.restype = ctypes.c_int
        L.b2o_bodies.argtypes = [vp, vp]
        L.b2o_pairs.restype = L_
        L.b2o_pairs.argtypes = [vp, vp, L_]
        L.b2o_query.restype = ctypes.c_int
        L.b2o_query.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int]
        _libs[precision] = L
    return _libs[precision]


def created_contacts(before, after):
    """(ids, shape_a, shape_b) of the contacts ``after`` has and ``before`` lacks,
    in the order the contact-id pool hands their ids out (its free list from the
    top, then fresh ids; S/id_pool.c:19-37)."""
    old = {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in before.live_contacts()}
    new = {}
    for c in after.live_contacts():
        key = (int(c["shape_a"]), int(c["shape_b"]))
        if old.get(int(c["id"])) != key:
            new[int(c["id"])] = key
    seq = [int(i) for i in before.free_ids[::-1]] + list(range(len(before.contacts), len(before.contacts) + len(new)))
    ids = [i for i in seq if i in new][:len(new)]
    a = [new[i][0] for i in ids]
    b = [new[i][1] for i in ids]
    return (np.array(ids, np.int32), np.array(a, np.int32), np.array(b, np.int32))


class Oracle:
    def __init__(self, spec: bytes, geoms: bytes, snapshot: bytes, precision: str = "f32", world: int = 0):
        self.L = _lib(precision)
        self._keep = (spec, geoms, snapshot)
        self.h = self.L.b2o_create(spec, len(spec), world, geoms, len(geoms), snapshot, len(snapshot))
        if not self.h:
            raise ValueError("oracle: malformed spec, geometry or snapshot")

    def close(self):
        if self.h:
            self.L.b2o_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def create_contacts(self, created):
        ids, a, b = (np.ascontiguousarray(x, np.int32) for x in created)
        rc = self.L.b2o_create_contacts(self.h, ids.ctypes.data, a.ctypes.data, b.ctypes.data, len(ids))
        if rc != 0:
            raise RuntimeError("oracle: created contact ids are not the id pool's allocations")

    def update_pairs(self) -> int:
        """The step's pair update over the trees: contacts created, ids from the pool."""
        n = self.L.b2o_update_pairs(self.h)
        if n < 0:
            raise RuntimeError("oracle: the snapshot holds no broad-phase trees (version < 3)")
        return n

    def collide(self):
        self.L.b2o_collide(self.h)

    def solve(self, dt: float = 1.0 / 60.0, substeps: int = 4):
        self.L.b2o_solve(self.h, dt, substeps)

    def export(self) -> bytes:
        n = self.L.b2o_export(self.h, None, 0)
        buf = ctypes.create_string_buffer(n)
        self.L.b2o_export(self.h, buf, n)
        return buf.raw

    def bodies(self) -> np.ndarray:
        n = len(self._body_probe())
        out = np.zeros((n, 7), np.float64)
        self.L.b2o_bodies(self.h, out.ctypes.data)
        return out

    def _body_probe(self):
        from forge2d_amd import snapshot as S

        return S.parse_one(self.export()).bodies

    def query(self, queries: np.ndarray, cap: int = 64):
        """f2dQuery records (forge2d_amd.make_queries) -> (counts, hits[n, cap]); explosions
        are applied to the oracle's world."""
        from forge2d_amd import HIT_DTYPE, QUERY_DTYPE

        queries = np.ascontiguousarray(queries, QUERY_DTYPE)
        n = len(queries)
        hits = np.zeros((max(n, 1), cap), HIT_DTYPE)
        counts = np.zeros(max(n, 1), np.int32)
        self.L.b2o_query(self.h, queries.ctypes.data, n, hits.ctypes.data, counts.ctypes.data, cap)
        return counts[:n], hits[:n]

    def pairs(self) -> np.ndarray:
        n = self.L.b2o_pairs(self.h, None, 0)
        out = np.zeros(max(n, 1), np.int64)
        self.L.b2o_pairs(self.h, out.ctypes.data, n)
        return np.sort(out[:n])


def step(spec: bytes, geoms: bytes, before_raw: bytes, before, after=None, precision: str = "f32",
         dt: float = 1.0 / 60.0, substeps: int = 4) -> "Oracle":
    """One full step from snapshot ``before``: the pair update over its trees when it
    has them (version 3), else the contacts created in that step as ``after`` records
    them; then collide and solve."""
    o = Oracle(spec, geoms, before_raw, precision)
    if getattr(before, "trees", None) is not None:
        o.update_pairs()
    elif after is not None:
        o.create_contacts(created_contacts(before, after))
    o.collide()
    o.solve(dt, substeps)
    return o
