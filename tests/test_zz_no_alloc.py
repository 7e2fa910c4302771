"""The allocation-free query entry point f2dg_query_device_static against the
library's device query entry point f2dg_query_device: both read the same device
arrays and must write the same counts and hit records bit for bit (the hit slots
beyond each count keep their sentinel bytes in both). The queries come shuffled
across worlds, so both group them by world in their scratch (the static array, or a
buffer the library allocates). Two calls in a row, the second on a smaller batch
with fewer queries, so stale grouping of the first call in the static array cannot
leak into the second. A batch that is not uploaded, or a call with more queries and
worlds than the static array holds, is refused with an error."""
import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes

pytestmark = pytest.mark.gpu

ALL = 0xFFFFFFFF
CAP = 64


def world_queries(b, worlds, per_world, seed):
    """Closest and every-hit rays and AABB overlaps over each world's bodies."""
    rng = np.random.default_rng(seed)
    rows = []
    for wi in range(worlds):
        bodies = b.bodies(wi)
        lo = bodies[:, :2].min(axis=0) - 2.0
        hi = bodies[:, :2].max(axis=0) + 2.0
        size = float(np.max(hi - lo))
        for k in range(per_world):
            o = rng.uniform(lo, hi)
            if k % 3 == 2:
                e = rng.uniform(0.05, 0.25, 2) * size
                rows.append((wi, F.Q_OVERLAP_AABB, [o[0] - e[0], o[1] - e[1], o[0] + e[0], o[1] + e[1]]))
            else:
                a = rng.uniform(-np.pi, np.pi)
                ln = rng.uniform(0.1, 1.0) * size
                kind = F.Q_RAY_CLOSEST if k % 3 == 0 else F.Q_RAY_ALL
                rows.append((wi, kind, [o[0], o[1], ln * np.cos(a), ln * np.sin(a)]))
    return F.make_queries(rows)


def run(entry, b, q):
    import torch

    dq = torch.from_numpy(q.view(np.uint8).copy()).cuda()
    dh = torch.full((len(q) * CAP * F.HIT_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device="cuda")
    dc = torch.full((len(q),), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = entry(b.h, dq.data_ptr(), len(q), dh.data_ptr(), dc.data_ptr(), CAP)
    assert rc == 0, F.lib().f2dg_last_error(b.h)
    b.synchronize()
    return dc.cpu().numpy(), dh.cpu().numpy()


def test_static_entry_point_matches_device_entry_point(gpu):
    L = F.lib()
    for worlds, steps, per_world, seed in ((96, 12, 24, 1), (24, 5, 9, 2)):
        print(f"[no-alloc] {worlds} worlds, {per_world} queries per world", flush=True)
        b = F.Batch(scenes.config4(worlds=worlds, lo=10, hi=300))
        b.step(steps=steps)
        assert b.sync() == 0
        q = world_queries(b, worlds, per_world, seed)
        q = q[np.random.default_rng(seed).permutation(len(q))]
        ref_counts, ref_hits = run(L.f2dg_query_device, b, q)
        counts, hits = run(L.f2dg_query_device_static, b, q)
        assert (ref_counts > 0).sum() > len(q) // 4  # the queries do hit shapes
        assert np.array_equal(counts, ref_counts)
        assert hits.tobytes() == ref_hits.tobytes()
        b.close()


def test_static_entry_point_refuses_what_it_cannot_hold(gpu):
    L = F.lib()
    b = F.Batch(scenes.config4(worlds=4, lo=10, hi=50))
    assert L.f2dg_query_device_static(b.h, None, 1, None, None, 1) == -1  # not uploaded: it allocates nothing
    assert b"upload" in L.f2dg_last_error(b.h)
    b.close()
    small = F.Batch(scenes.config4(worlds=4, lo=10, hi=50))
    small.upload()
    assert L.f2dg_query_device_static(small.h, None, 1 << 20, None, None, 1) == -1
    assert b"1048576" in L.f2dg_last_error(small.h)
    small.close()
