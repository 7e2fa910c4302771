"""Serving loop (continuous batching) on the device.

* A world's result is the same alone and in the mix: after a burst of requests
  over a small population, every resident world is compared bit for bit with the
  same world stepped alone in its own batch for the same number of steps.
* Admission respects the HBM budget: with a budget that holds only a few worlds,
  idle batches are evicted, the device bytes in use never exceed the budget and
  every request still completes.
* Eviction keeps a world's state: every world has run exactly the steps its
  requests asked for, across evictions and re-admissions, and a world evicted and
  re-admitted equals the same world stepped alone that many steps.
* A request for a world larger than the whole budget is refused and the run ends.
"""
import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import serving
from forge2d_amd import snapshot as S

pytestmark = pytest.mark.gpu


def small_population():
    return serving.Population(12, lo=40, hi=400, templates=6, seed=3)


def test_result_same_alone_and_in_mix(gpu):
    pop = small_population()
    reqs = serving.poisson_requests(150.0, 0.6, pop, seed=4, steps_choices=(1, 4, 9))
    eng = serving.ServingEngine(pop, hbm_budget=8 << 30)
    print(f"[serving] {len(reqs)} requests over {len(pop.sizes)} worlds", flush=True)
    stats = eng.run(reqs, horizon=0.6)
    assert stats["completed"] == len(reqs)
    checked = 0
    for w, res in sorted(eng.resident.items()):
        mine = S.parse_one(res.batch.batch.export(res.index))
        alone = F.Batch(b"".join([serving.scenes.MAGIC, np.array([1, 0], "<i4").tobytes(), pop.spec(w)]))
        alone.step(steps=eng.total_steps[w])
        assert alone.sync() == 0
        ref = S.parse_one(alone.export(0))
        alone.close()
        for f in ("px", "py", "qc", "qs", "vx", "vy", "w"):
            assert np.array_equal(mine.bodies[f], ref.bodies[f]), (w, f)
        assert np.array_equal(mine.pair_keys(), ref.pair_keys()), w
        checked += 1
    eng.close()
    assert checked >= 6


def eng_estimate(pop, world):
    return serving.ServingEngine(pop, 0, use_streams=False)._estimate(world)


def test_admission_respects_budget(gpu):
    pop = small_population()
    reqs = serving.poisson_requests(100.0, 0.5, pop, seed=9, steps_choices=(1, 3))
    big = int(np.argmax(pop.sizes))
    eng = serving.ServingEngine(pop, hbm_budget=3 * eng_estimate(pop, big))
    stats = eng.run(reqs, horizon=0.5)
    assert stats["completed"] == len(reqs)
    assert eng.used <= eng.budget
    assert stats["evicted_batches"] > 0
    asked = {}
    for r in reqs:
        asked[r.world] = asked.get(r.world, 0) + r.steps
    readmitted = 0
    for w, n in asked.items():
        assert eng.total_steps[w] == n, w
        assert eng.world_step(w) == n, w
        res = eng.resident.get(w)
        if res is not None and res.steps_done < n:
            # evicted earlier and continued from its exported state
            mine = S.parse_one(res.batch.batch.export(res.index))
            alone = F.Batch(b"".join([serving.scenes.MAGIC, np.array([1, 0], "<i4").tobytes(), pop.spec(w)]))
            alone.step(steps=n)
            assert alone.sync() == 0
            ref = S.parse_one(alone.export(0))
            alone.close()
            for f in ("px", "py", "qc", "qs", "vx", "vy", "w"):
                assert np.array_equal(mine.bodies[f], ref.bodies[f]), (w, f)
            readmitted += 1
    print(f"[serving] {readmitted} worlds continued after eviction", flush=True)
    eng.close()


def test_evicted_world_continues(gpu):
    pop = small_population()
    big = int(np.argmax(pop.sizes))
    small = int(np.argmin(pop.sizes))
    eng = serving.ServingEngine(pop, hbm_budget=eng_estimate(pop, big))
    # the big world evicts the small one, which then comes back for 3 more steps
    reqs = [serving.Request(0, small, 2, 0.0), serving.Request(1, big, 1, 0.2), serving.Request(2, small, 3, 0.4)]
    stats = eng.run(reqs, horizon=0.5)
    assert stats["completed"] == 3 and stats["evicted_batches"] >= 1
    assert eng.total_steps[small] == 5 and eng.world_step(small) == 5
    res = eng.resident[small]
    assert res.steps_done == 3
    mine = S.parse_one(res.batch.batch.export(res.index))
    alone = F.Batch(b"".join([serving.scenes.MAGIC, np.array([1, 0], "<i4").tobytes(), pop.spec(small)]))
    alone.step(steps=5)
    assert alone.sync() == 0
    ref = S.parse_one(alone.export(0))
    alone.close()
    for f in ("px", "py", "qc", "qs", "vx", "vy", "w"):
        assert np.array_equal(mine.bodies[f], ref.bodies[f]), f
    assert np.array_equal(mine.pair_keys(), ref.pair_keys())
    eng.close()


def test_oversize_world_is_refused(gpu):
    pop = small_population()
    big = int(np.argmax(pop.sizes))
    small = int(np.argmin(pop.sizes))
    eng = serving.ServingEngine(pop, hbm_budget=eng_estimate(pop, big) - 1)
    reqs = [serving.Request(0, big, 2, 0.0), serving.Request(1, small, 2, 0.0)]
    stats = eng.run(reqs, horizon=0.1)
    assert stats["refused"] == 1 and stats["completed"] == 1
    eng.close()


def test_admission_control_bounds_queued_steps(gpu):
    pop = small_population()
    reqs = serving.poisson_requests(400.0, 0.5, pop, seed=6, steps_choices=(8, 60))
    eng = serving.ServingEngine(pop, hbm_budget=8 << 30, max_backlog_s=0.05)
    assert eng.preload()
    stats = eng.run(reqs, horizon=0.5)
    # every request is either refused on arrival or completed; a world never owes
    # more than the bound allows at the iteration time measured when it was admitted
    assert stats["refused"] > 0
    assert stats["completed"] + stats["refused"] == len(reqs)
    asked = {}
    for r in reqs:
        if r.done >= 0:
            asked[r.world] = asked.get(r.world, 0) + r.steps
    for w, n in asked.items():
        assert eng.total_steps[w] == n, w
    eng.close()
