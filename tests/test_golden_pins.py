"""Pins of the golden fixtures (the reference's own outputs) against what the
mathematics fixes, independently of both engines:

* free fall: before the first impact every body's velocity and position follow
  the substep integrator v += h*g, dp += h*v (S/solver.c:65-198) evaluated in
  float32 — and agree with the exact float64 closed forms v = n*h*g,
  y = y0 + h^2*g*n(n+1)/2 to float32 rounding;
* soft contact constants: contactSpeed = maxPushSpeed / massScale(2*hertz) with
  b2MakeSoft's closed form (S/solver.h:140-182) in float64;
* mass: a unit box of density 1 has invMass 1 and invI 6 (I = m(w^2+h^2)/12); a
  circle of radius r has mass pi r^2 (S/geometry.c:220-399);
* broadphase: the live contact set after step k equals the brute-force set of
  overlapping fat AABBs of step k-1 (minus same-body, static-static and
  joint-connected pairs) — the invariant the device broadphase is built on;
* circle-circle manifolds: one point, id 0, normal = unit centre difference,
  separation = distance - rA - rB (textbook, S/manifold.c:40-74);
* the reference's colouring is valid: no dynamic body twice in a colour, no
  static contact in colour 0 (S/constraint_graph.c:68-135).
"""
import itertools
import math

import numpy as np
import pytest

from forge2d_amd import scenes
from tests import parity_util as P
from tests.conftest import FIXTURE_NAMES, load_fixture

F32 = np.float32


def spec_world(name):
    spec, _, _ = load_fixture(name)
    return scenes.parse(spec)[0]


def test_free_fall_integrator():
    spec, snaps, _ = load_fixture("bounce")
    h_, bodies, shapes, _ = spec_world("bounce")
    s40 = next(s for s in snaps if s.step == 40)
    dt = F32(1.0 / 60.0)
    h = F32(dt / F32(4))
    g = F32(-10.0)
    n = 40 * 4
    checked = 0
    for i in range(len(bodies)):
        if bodies[i]["type"] != 2:
            continue
        v = F32(0.0)
        y = F32(snaps[0].bodies[i]["cy"])
        for step in range(40):
            dp = F32(0.0)
            for sub in range(4):
                v = F32(F32(h * g) + F32(F32(1.0) * v))
                dp = F32(dp + F32(h * v))
            y = F32(y + dp)
        # only bodies whose free-fall path keeps them clear of the ground (top at y=0)
        # by more than the fat-AABB margins have never touched anything
        if float(y) - float(s40.bodies[i]["max_ext"]) < 0.1:
            continue
        assert s40.bodies[i]["vy"] == v
        assert s40.bodies[i]["cy"] == y
        hd = 1.0 / 240.0
        assert abs(float(v) - n * hd * -10.0) < 1e-4
        y_exact = float(snaps[0].bodies[i]["cy"]) + hd * hd * -10.0 * n * (n + 1) / 2
        assert abs(float(y) - y_exact) < 1e-4
        checked += 1
    assert checked >= 3


def make_soft64(hertz, zeta, h):
    omega = 2.0 * math.pi * hertz
    a1 = 2.0 * zeta + h * omega
    a2 = h * omega * a1
    a3 = 1.0 / (1.0 + a2)
    return omega / a1, a2 * a3, a3


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_contact_speed_closed_form(name):
    _, snaps, _ = load_fixture(name)
    s = snaps[-1]
    h = 1.0 / 240.0
    hertz = min(30.0, 0.125 * 240.0)
    bias, mass_scale, impulse_scale = make_soft64(2.0 * hertz, 10.0, h)
    assert abs(mass_scale + impulse_scale - 1.0) < 1e-12
    assert abs(float(s.header["contact_speed"]) - 3.0 / mass_scale) < 1e-5 * (3.0 / mass_scale)
    assert abs(float(s.header["inv_h"]) - 240.0) < 1e-3


def test_box_and_circle_mass():
    _, snaps, _ = load_fixture("pyramid12")
    b = snaps[0].bodies[1:]
    assert np.all(np.abs(b["inv_mass"] - 1.0) < 1e-6)
    assert np.all(np.abs(b["inv_i"] - 6.0) < 1e-5)
    h_, bodies, shapes, _ = spec_world("mixed")
    s0 = load_fixture("mixed")[1][0]
    for si, sh in enumerate(shapes):
        if sh["kind"] == scenes.K_CIRCLE:
            bi = int(sh["body"])
            m = math.pi * float(sh["radius"]) ** 2 * float(sh["density"])
            assert abs(1.0 / float(s0.bodies[bi]["inv_mass"]) - m) < 1e-5 * m


def brute_force_pairs(prev, spec_bodies, joints):
    fat = prev.shapes["fat"].astype(np.float64)
    body = prev.shapes["body"]
    types = spec_bodies["type"]
    no_collide = set()
    for j in joints:
        if not j["collide_connected"]:
            no_collide.add((int(j["body_a"]), int(j["body_b"])))
            no_collide.add((int(j["body_b"]), int(j["body_a"])))
    out = set()
    n = len(fat)
    for a, b in itertools.combinations(range(n), 2):
        ba, bb = int(body[a]), int(body[b])
        if ba == bb or (types[ba] != 2 and types[bb] != 2) or (ba, bb) in no_collide:
            continue
        fa, fb = fat[a], fat[b]
        if fb[0] > fa[2] or fb[1] > fa[3] or fa[0] > fb[2] or fa[1] > fb[3]:
            continue
        out.add((min(a, b), max(a, b)))
    return out


@pytest.mark.parametrize("name", ["mixed", "bounce", "overflow", "pyramid12"])
def test_contact_set_is_fat_aabb_overlap_set(name):
    spec, snaps, _ = load_fixture(name)
    h_, bodies, shapes, joints = scenes.parse(spec)[0]
    checked = 0
    for i in range(1, len(snaps)):
        if snaps[i].step != snaps[i - 1].step + 1:
            continue
        expect = brute_force_pairs(snaps[i - 1], bodies, joints)
        got = {(min(int(a), int(b)), max(int(a), int(b))) for a, b in
               zip(snaps[i].live_contacts()["shape_a"], snaps[i].live_contacts()["shape_b"])}
        # segment-segment pairs have no manifold function and never become contacts
        seg = {k for k in expect if shapes[k[0]]["kind"] == scenes.K_SEGMENT and shapes[k[1]]["kind"] == scenes.K_SEGMENT}
        assert got == expect - seg, (name, snaps[i].step)
        checked += 1
    assert checked > 0


def test_circle_manifolds_textbook():
    """Manifolds of step k are computed from the poses after step k-1: for two
    circles centred on their bodies the normal is the unit centre difference and
    the separation is |cB - cA| - rA - rB."""
    spec, snaps, _ = load_fixture("mixed")
    h_, bodies, shapes, _ = scenes.parse(spec)[0]
    checked = 0
    for i in range(1, len(snaps)):
        prev, s = snaps[i - 1], snaps[i]
        if s.step != prev.step + 1:
            continue
        for c in s.live_contacts():
            a, b = int(c["shape_a"]), int(c["shape_b"])
            if shapes[a]["kind"] != scenes.K_CIRCLE or shapes[b]["kind"] != scenes.K_CIRCLE:
                continue
            if int(c["point_count"]) == 0:
                continue
            ba, bb = prev.bodies[int(shapes[a]["body"])], prev.bodies[int(shapes[b]["body"])]
            d = np.array([float(bb["px"]) - float(ba["px"]), float(bb["py"]) - float(ba["py"])])
            dist = float(np.linalg.norm(d))
            n = np.array([float(c["nx"]), float(c["ny"])])
            assert np.allclose(n, d / dist, atol=1e-5)
            sep = dist - float(shapes[a]["radius"]) - float(shapes[b]["radius"])
            assert abs(float(c["pts"][0]["sep"]) - sep) < 1e-5
            assert int(c["pts"][0]["id"]) == 0 and int(c["point_count"]) == 1
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_reference_colouring_is_valid(name):
    _, snaps, _ = load_fixture(name)
    for s in snaps:
        assert P.coloring_violations(s) == 0
