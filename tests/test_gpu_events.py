"""Contact and body events (S/world.c:587-690, S/contact.c:354-372 begin/end;
S/solver.c:1758-1817 hits; S/solver.c:605-610 body moves) against the
reference's own events, recorded by the fixture harness for every step of the
"events" fixture (bouncing circles, capsules and boxes with contact and hit events
enabled on every shape), and of the "sleep" fixture, whose body move events carry
the order of the awake set and the fellAsleep flag of islands put to sleep
(S/solver_set.c:196-205).

For every step k: import the reference state after k-1, step once, read the
events through the C-ABI getters (the flat shim's record layouts) and compare
with the reference's events of step k. Contact ids equal the reference's, and begin
and end events must be equal record for record and in order. Hit events are compared
as sets (the reference orders them by its colour arrays). Body move events must equal
the reference's transforms.
"""
import os
import struct

import numpy as np
import pytest

import forge2d_amd as F
from tests import parity_util as P
from tests.conftest import GOLDEN, load_fixture

pytestmark = pytest.mark.gpu


def reference_events(name):
    raw = np.load(os.path.join(GOLDEN, name + ".npz"))["events"].tobytes()
    out, o = {}, 0
    while o < len(raw):
        step, nb, ne, nh, nm = struct.unpack_from("<5i", raw, o)
        o += 20
        recs = {}
        for key, n, width in (("begin", nb, 13), ("end", ne, 4), ("hit", nh, 9), ("move", nm, 7)):
            recs[key] = np.frombuffer(raw, "<f8", n * width, o).reshape(n, width)
            o += 8 * n * width
        out[step] = recs
    return out


def sorted_rows(a):
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.mark.parametrize("name", ["events", "sleep"])
def test_contact_and_body_events_match_reference(gpu, name):
    spec, snaps, raws = load_fixture(name)
    ref_events = reference_events(name)
    checked = 0
    for i in range(1, len(snaps)):
        prev, ref = snaps[i - 1], snaps[i]
        k = ref.step
        if k not in ref_events or prev.step != k - 1:
            continue
        b = F.Batch(spec)
        b.import_state(0, raws[i - 1])
        b.step(steps=1)
        assert b.sync() == 0
        mine = b.snapshot(0)
        ev = b.contact_events(0)
        re = ref_events[k]
        assert {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in mine.live_contacts()} == \
               {int(c["id"]): (int(c["shape_a"]), int(c["shape_b"])) for c in ref.live_contacts()}, k
        for key in ("begin", "end"):
            assert np.array_equal(ev[key], re[key]), (k, key)
        assert np.array_equal(ev["hit"], re["hit"]), k
        assert P.color_map(mine) == P.color_map(ref), k
        mv = b.body_events(0)
        assert np.array_equal(mv[:, [0, 2, 3, 4, 5, 6]], re["move"][:, [0, 2, 3, 4, 5, 6]]), k
        checked += 1
        b.close()
    assert checked >= 100


def test_world_events_through_flat_abi(gpu):
    """A ball dropped on a ground box through the f2d_* ABI: a begin-touch event
    when it lands, hit events above the threshold, a move event every step."""
    import ctypes

    from tests.test_gpu_api import body, shape_def, world

    L = F.lib()
    w = world()
    ground = body(w, 0, 0.0, -1.0)
    d = shape_def()
    d.restitution = 0.5
    d.enable_contact_events = 1
    d.enable_hit_events = 1
    out = (ctypes.c_int32 * 2)()
    L.f2d_create_box_shape(ground[0], ground[1], ctypes.byref(d), 10.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, out)
    ball = body(w, 2, 0.0, 2.0)
    L.f2d_create_circle_shape(ball[0], ball[1], ctypes.byref(d), 0.0, 0.0, 0.5, out)
    begins = hits = 0
    pd = ctypes.POINTER(ctypes.c_double)
    for _ in range(90):
        L.f2d_world_step(w, 1.0 / 60.0, 4)
        counts = (ctypes.c_int32 * 3)()
        L.f2d_world_get_contact_event_counts(w, counts)
        begins += counts[0]
        if counts[2]:
            rec = np.zeros(9 * counts[2], np.float64)
            assert L.f2d_world_get_contact_hit_events(w, rec.ctypes.data_as(pd)) == counts[2]
            assert rec[8] > 1.0
            hits += counts[2]
        assert L.f2d_world_get_body_event_count(w) == 1
    assert begins >= 1 and hits >= 1
    L.f2d_destroy_world(w)
