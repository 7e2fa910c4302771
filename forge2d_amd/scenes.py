"""Seeded scene generators and the binary scene-spec format.

A scene spec is a list of worlds. Each world holds body, shape and revolute-joint
records that map one-to-one onto the forge2d body/shape/joint definitions
(``F/lib/src/api/defs.dart:82-229``) and onto the flat create calls of the
reference's own C shim (``F/native/wasm/f2d_shim.c:123-139, 192-218, 459-515,
816-843``). Both the reference fixture harness (``oracle/ref_harness.c``) and
this package read the same file, so the two engines build bit-identical
scenes.

File layout (little-endian)::

    char[8] magic "F2DSCN01"; int32 world_count; int32 reserved
    per world: WORLD_DT header, body_count x BODY_DT, shape_count x SHAPE_DT,
               joint_count x JOINT_DT

Randomness is splitmix64 (Steele et al., 2014) with a fixed draw order per
world; the stream seed is derived from the config index and world index.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = b"F2DSCN01"

# body types (I/types.h b2BodyType)
STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
# shape kinds in the spec
K_CIRCLE, K_CAPSULE, K_SEGMENT, K_POLY, K_BOX = 0, 1, 2, 3, 4

# body flags
BF_BULLET, BF_FIXED_ROT, BF_FAST_ROT, BF_ENABLE_SLEEP, BF_AWAKE = 1, 2, 4, 8, 16
# shape flags
SF_CONTACT_EVENTS, SF_HIT_EVENTS, SF_INVOKE_CONTACT, SF_PRESOLVE, SF_UPDATE_MASS, SF_SENSOR = 1, 2, 4, 8, 16, 32
# world flags
WF_SLEEP, WF_CONTINUOUS, WF_WARM = 1, 2, 4

WORLD_DT = np.dtype([
    ("body_count", "<i4"), ("shape_count", "<i4"), ("joint_count", "<i4"), ("flags", "<i4"),
    ("gx", "<f4"), ("gy", "<f4"), ("restitution_threshold", "<f4"), ("hit_event_threshold", "<f4"),
    ("contact_hertz", "<f4"), ("contact_damping_ratio", "<f4"), ("max_contact_push_speed", "<f4"),
    ("maximum_linear_speed", "<f4"),
])
BODY_DT = np.dtype([
    ("type", "<i4"), ("px", "<f4"), ("py", "<f4"), ("qc", "<f4"), ("qs", "<f4"),
    ("vx", "<f4"), ("vy", "<f4"), ("w", "<f4"), ("lin_damp", "<f4"), ("ang_damp", "<f4"),
    ("gravity_scale", "<f4"), ("sleep_threshold", "<f4"), ("flags", "<i4"),
])
SHAPE_DT = np.dtype([
    ("body", "<i4"), ("kind", "<i4"), ("friction", "<f4"), ("restitution", "<f4"), ("rolling", "<f4"),
    ("tangent_speed", "<f4"), ("density", "<f4"), ("cat_lo", "<u4"), ("cat_hi", "<u4"),
    ("mask_lo", "<u4"), ("mask_hi", "<u4"), ("group", "<i4"), ("flags", "<i4"), ("npts", "<i4"),
    ("radius", "<f4"), ("data", "<f4", (16,)),
])
JOINT_DT = np.dtype([
    ("body_a", "<i4"), ("body_b", "<i4"), ("lax", "<f4"), ("lay", "<f4"), ("lbx", "<f4"), ("lby", "<f4"),
    ("ref_angle", "<f4"), ("enable_limit", "<i4"), ("lower", "<f4"), ("upper", "<f4"),
    ("enable_motor", "<i4"), ("motor_speed", "<f4"), ("max_motor_torque", "<f4"),
    ("enable_spring", "<i4"), ("hertz", "<f4"), ("damping", "<f4"), ("collide_connected", "<i4"),
    ("draw_size", "<f4"),
])
assert WORLD_DT.itemsize == 48 and BODY_DT.itemsize == 52
assert SHAPE_DT.itemsize == 124 and JOINT_DT.itemsize == 72

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 stream; ``uniform`` draws in [lo, hi) from the top 24 bits."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        u = (self.next() >> 40) / float(1 << 24)
        return float(np.float32(lo + (hi - lo) * u))

    def randint(self, lo: int, hi: int) -> int:
        """integer in [lo, hi] inclusive"""
        return lo + self.next() % (hi - lo + 1)


def stream_seed(config: int, world: int) -> int:
    return ((config + 1) * 0x100000001B3 + world * 0x9E3779B97F4A7C15) & MASK64


@dataclass
class World:
    header: np.ndarray
    bodies: list = field(default_factory=list)
    shapes: list = field(default_factory=list)
    joints: list = field(default_factory=list)

    def add_body(self, type_, px, py, angle=0.0, vx=0.0, vy=0.0, w=0.0, flags=BF_ENABLE_SLEEP | BF_AWAKE,
                 lin_damp=0.0, ang_damp=0.0, gravity_scale=1.0, sleep_threshold=0.05):
        qc, qs = rot_of(angle)
        self.bodies.append((type_, px, py, qc, qs, vx, vy, w, lin_damp, ang_damp, gravity_scale, sleep_threshold,
                            flags))
        return len(self.bodies) - 1

    def add_shape(self, body, kind, data, npts=0, radius=0.0, friction=0.6, restitution=0.0, density=1.0,
                  flags=SF_INVOKE_CONTACT | SF_UPDATE_MASS, group=0, rolling=0.0, tangent_speed=0.0,
                  category=(1, 0), mask=(0xFFFFFFFF, 0xFFFFFFFF)):
        d = list(data) + [0.0] * (16 - len(data))
        self.shapes.append((body, kind, friction, restitution, rolling, tangent_speed, density, category[0],
                            category[1], mask[0], mask[1], group, flags, npts, radius, d))
        return len(self.shapes) - 1

    def add_box(self, body, hx, hy, cx=0.0, cy=0.0, angle=0.0, radius=0.0, **kw):
        c, s = rot_of(angle)
        return self.add_shape(body, K_BOX, [hx, hy, cx, cy, c, s], radius=radius, **kw)

    def add_circle(self, body, r, cx=0.0, cy=0.0, **kw):
        return self.add_shape(body, K_CIRCLE, [cx, cy], radius=r, **kw)

    def add_capsule(self, body, c1, c2, r, **kw):
        return self.add_shape(body, K_CAPSULE, [c1[0], c1[1], c2[0], c2[1]], radius=r, **kw)

    def add_segment(self, body, p1, p2, **kw):
        return self.add_shape(body, K_SEGMENT, [p1[0], p1[1], p2[0], p2[1]], **kw)

    def add_polygon(self, body, pts, radius=0.0, **kw):
        flat = [c for p in pts for c in p]
        return self.add_shape(body, K_POLY, flat, npts=len(pts), radius=radius, **kw)

    def add_revolute(self, a, b, la, lb, ref_angle=0.0, enable_limit=False, lower=0.0, upper=0.0,
                     enable_motor=False, motor_speed=0.0, max_motor_torque=0.0, enable_spring=False,
                     hertz=0.0, damping=0.0, collide_connected=False):
        self.joints.append((a, b, la[0], la[1], lb[0], lb[1], ref_angle, int(enable_limit), lower, upper,
                            int(enable_motor), motor_speed, max_motor_torque, int(enable_spring), hertz, damping,
                            int(collide_connected), 0.25))
        return len(self.joints) - 1

    def arrays(self):
        b = np.array(self.bodies, BODY_DT) if self.bodies else np.zeros(0, BODY_DT)
        s = np.array(self.shapes, SHAPE_DT) if self.shapes else np.zeros(0, SHAPE_DT)
        j = np.array(self.joints, JOINT_DT) if self.joints else np.zeros(0, JOINT_DT)
        h = self.header.copy()
        h["body_count"], h["shape_count"], h["joint_count"] = len(b), len(s), len(j)
        return h, b, s, j


def rot_of(angle: float):
    """b2MakeRot for the angles used here is only called with 0 or with the
    spec's own cos/sin; the spec stores (cos, sin) as float32 of the double
    values so both engines receive identical rotations."""
    return float(np.float32(math.cos(angle))), float(np.float32(math.sin(angle)))


def world_header(sleep=False, continuous=True, warm=True, gravity=(0.0, -10.0)):
    h = np.zeros((), WORLD_DT)
    flags = (WF_SLEEP if sleep else 0) | (WF_CONTINUOUS if continuous else 0) | (WF_WARM if warm else 0)
    h["flags"] = flags
    h["gx"], h["gy"] = gravity
    # b2DefaultWorldDef, S/types.c:9-27
    h["restitution_threshold"] = 1.0
    h["hit_event_threshold"] = 1.0
    h["contact_hertz"] = 30.0
    h["contact_damping_ratio"] = 10.0
    h["max_contact_push_speed"] = 3.0
    h["maximum_linear_speed"] = 400.0
    return h


# ----------------------------------------------------------------------------- configs

def pyramid_world(rows: int, sleep=False, continuous=True, box_half=0.5, spacing=1.0, density=1.0,
                  friction=0.6, restitution=0.0, ground_half=(50.0, 1.0), ground_y=-1.0) -> World:
    """Config 1 / config 3 scene: static ground box plus a triangular pyramid.
    Row r holds rows-r boxes at x_i = (i - (rows-r-1)/2) * spacing, y = 0.5 + r * spacing."""
    w = World(world_header(sleep=sleep, continuous=continuous))
    g = w.add_body(STATIC, 0.0, ground_y)
    w.add_box(g, ground_half[0], ground_half[1], friction=friction, restitution=restitution)
    for r in range(rows):
        n = rows - r
        for i in range(n):
            x = (i - (n - 1) / 2.0) * spacing
            y = box_half + r * spacing
            b = w.add_body(DYNAMIC, x, y)
            w.add_box(b, box_half, box_half, density=density, friction=friction, restitution=restitution)
    return w


def _random_shape(w: World, rng: SplitMix64, body: int):
    """Config 2 shape mix: 40% circles, 30% capsules, 30% 3-8 vertex polygons."""
    friction = rng.uniform(0.2, 0.8)
    restitution = rng.uniform(0.0, 0.3)
    u = rng.uniform(0.0, 1.0)
    if u < 0.4:
        r = rng.uniform(0.2, 0.6)
        w.add_circle(body, r, friction=friction, restitution=restitution)
    elif u < 0.7:
        length = rng.uniform(0.4, 1.2)
        r = rng.uniform(0.15, 0.3)
        h = 0.5 * length
        w.add_capsule(body, (-h, 0.0), (h, 0.0), r, friction=friction, restitution=restitution)
    else:
        n = rng.randint(3, 8)
        rad = rng.uniform(0.3, 0.8)
        pts = []
        for k in range(n):
            jitter = rng.uniform(-0.3, 0.3) * (math.pi / n)
            a = 2.0 * math.pi * k / n + jitter
            pts.append((float(np.float32(rad * math.cos(a))), float(np.float32(rad * math.sin(a)))))
        w.add_polygon(body, pts, friction=friction, restitution=restitution)


def container_world(n_bodies: int, half_width: float, height: float, seed: int, sleep=False,
                    continuous=True, chains: int = 0, drop_band=(0.5, 1.0)) -> World:
    """Static segment container [-hw, hw] x [0, height] with n seeded bodies in the
    upper band of the box, plus optional revolute chains hanging from the top."""
    rng = SplitMix64(seed)
    w = World(world_header(sleep=sleep, continuous=continuous))
    g = w.add_body(STATIC, 0.0, 0.0)
    hw = half_width
    for p1, p2 in (((-hw, 0.0), (hw, 0.0)), ((-hw, 0.0), (-hw, height)), ((hw, 0.0), (hw, height)),
                   ((-hw, height), (hw, height))):
        w.add_segment(g, p1, p2)
    # revolute chains of 5-12 links, link box 0.1 x 0.4 (full size), limits +-0.5 rad
    for c in range(chains):
        links = rng.randint(5, 12)
        ax = rng.uniform(-0.8 * hw, 0.8 * hw)
        ay = height - 1.0
        prev = g
        prev_anchor = (ax, ay)
        for k in range(links):
            y = ay - 0.2 - 0.4 * k
            b = w.add_body(DYNAMIC, ax, y)
            w.add_box(b, 0.05, 0.2, friction=0.6)
            if prev == g:
                w.add_revolute(g, b, prev_anchor, (0.0, 0.2), enable_limit=True, lower=-0.5, upper=0.5)
            else:
                w.add_revolute(prev, b, (0.0, -0.2), (0.0, 0.2), enable_limit=True, lower=-0.5, upper=0.5)
            prev = b
    y_lo = drop_band[0] * height
    y_hi = drop_band[1] * height - 1.0
    for i in range(n_bodies):
        x = rng.uniform(-hw + 1.0, hw - 1.0)
        y = rng.uniform(y_lo, y_hi)
        angle = rng.uniform(-math.pi, math.pi)
        b = w.add_body(DYNAMIC, x, y, angle=angle)
        _random_shape(w, rng, b)
    return w


def config1(sleep=False) -> list[World]:
    return [pyramid_world(40, sleep=sleep)]


def config2(sleep=False, n=20000) -> list[World]:
    return [container_world(n, 100.0, 200.0, stream_seed(2, 0), sleep=sleep)]


def config3(worlds=4096, sleep=False, rows_lo=5, rows_hi=30) -> list[World]:
    out = []
    for i in range(worlds):
        rng = SplitMix64(stream_seed(3, i))
        out.append(pyramid_world(rng.randint(rows_lo, rows_hi), sleep=sleep))
    return out


def config4(worlds=16384, sleep=False, lo=10, hi=500) -> list[World]:
    out = []
    for i in range(worlds):
        rng = SplitMix64(stream_seed(4, i))
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        chains = rng.randint(0, 8)
        # keep the total dynamic body count at n: chains take their links from the budget
        hw = max(8.0, math.sqrt(n) * 1.5)
        w = container_world(max(n - 8 * chains, n // 2), hw, 2.0 * hw, stream_seed(40, i), sleep=sleep,
                            chains=chains if n >= 40 else 0)
        out.append(w)
    return out


# ----------------------------------------------------------------------------- io

def to_bytes(worlds: list[World]) -> bytes:
    parts = [MAGIC, struct.pack("<ii", len(worlds), 0)]
    for w in worlds:
        h, b, s, j = w.arrays()
        parts += [h.tobytes(), b.tobytes(), s.tobytes(), j.tobytes()]
    return b"".join(parts)


def write(path: str, worlds: list[World]):
    with open(path, "wb") as f:
        f.write(to_bytes(worlds))


def parse(buf: bytes):
    """Inverse of ``to_bytes``: list of (header, bodies, shapes, joints) arrays."""
    assert buf[:8] == MAGIC, "not a scene spec"
    (n, _) = struct.unpack_from("<ii", buf, 8)
    off = 16
    out = []
    for _ in range(n):
        h = np.frombuffer(buf, WORLD_DT, 1, off)[0]
        off += WORLD_DT.itemsize
        b = np.frombuffer(buf, BODY_DT, int(h["body_count"]), off)
        off += b.nbytes
        s = np.frombuffer(buf, SHAPE_DT, int(h["shape_count"]), off)
        off += s.nbytes
        j = np.frombuffer(buf, JOINT_DT, int(h["joint_count"]), off)
        off += j.nbytes
        out.append((h, b, s, j))
    return out
