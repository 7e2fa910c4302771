"""Serving loop (continuous batching) on the device.

* A world's result is the same alone and in the mix: after a burst of requests
  over a small population, every resident world is compared bit for bit with the
  same world stepped alone in its own batch for the same number of steps.
* Admission respects the HBM budget: with a budget that holds only a few worlds,
  idle batches are evicted, the device bytes in use never exceed the budget and
  every request still completes.
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
        alone.step(steps=res.steps_done)
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
    eng.close()
