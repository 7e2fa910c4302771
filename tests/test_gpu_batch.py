"""GPU tests of batching and trajectories.

* a world stepped inside a batch of many worlds gives the same bits as the same
  world stepped alone (one launch per world), so batching never changes results;
* stepping a subset of worlds (serving) leaves the others untouched;
* every workgroup size the autotuner may pick (64 to 1024 lanes per world) gives
  the same bits;
* trajectory invariants over 256 steps at full configuration shape: a settled
  pyramid keeps its height, nothing tunnels through the ground, kinetic energy
  decays (config 1 and a config-3 batch).
"""
import numpy as np
import pytest

import forge2d_amd as F
from forge2d_amd import scenes

pytestmark = pytest.mark.gpu


def test_batched_equals_one_at_a_time(gpu):
    worlds = scenes.config3(worlds=24, rows_lo=5, rows_hi=14)
    a = F.Batch(worlds)
    b = F.Batch(worlds)
    a.step(steps=12)
    b.step_each(steps=12)
    assert a.sync() == 0 and b.sync() == 0
    for w in range(len(worlds)):
        assert np.array_equal(a.bodies(w), b.bodies(w)), w


def test_mixed_batch_equals_alone(gpu):
    worlds = scenes.config4(worlds=12, lo=10, hi=120)
    batch = F.Batch(worlds)
    batch.step(steps=8)
    assert batch.sync() == 0
    for w in (0, 5, 11):
        alone = F.Batch([worlds[w]])
        alone.step(steps=8)
        assert alone.sync() == 0
        assert np.array_equal(batch.bodies(w), alone.bodies(0)), w


@pytest.mark.parametrize("block", [64, 128, 512, 1024])
def test_workgroup_size_does_not_change_results(gpu, block):
    worlds = scenes.config4(worlds=6, lo=40, hi=400) + [scenes.container_world(1500, 30.0, 50.0, 7005, chains=3)]
    ref = F.Batch(worlds)
    ref.upload()
    ref.set_block(256)
    other = F.Batch(worlds)
    other.upload()
    other.set_block(block)
    assert {c["lanes"] for c in other.size_classes()} == {block}
    print(f"[batch] workgroup {block} lanes vs 256", flush=True)
    ref.step(steps=16)
    other.step(steps=16)
    assert ref.sync() == 0 and other.sync() == 0
    for w in range(len(worlds)):
        assert ref.export(w) == other.export(w), w


def test_autotune_advances_every_world_equally(gpu):
    """The autotuner's trial steps advance worlds of small classes (3 lane
    candidates) and of large classes (5 candidates) by the same step count."""
    worlds = scenes.config4(worlds=4, lo=10, hi=60) + [scenes.container_world(600, 20.0, 40.0, 7006, chains=2)]
    b = F.Batch(worlds)
    b.upload()
    assert min(c["limit"] for c in b.size_classes()) <= 256 < max(c["limit"] for c in b.size_classes())
    b.autotune(steps=2)
    assert b.sync() == 0
    steps = {b.counters(w)["step"] for w in range(len(worlds))}
    assert steps == {18}, steps


def test_online_tuning_does_not_change_results(gpu):
    """Online lane tuning (trial rounds every 5 steps, classes stepped one after
    another at each candidate size) gives the same bits as plain stepping."""
    worlds = scenes.config4(worlds=8, lo=20, hi=400) + [scenes.container_world(1500, 30.0, 50.0, 7007, chains=3)]
    ref = F.Batch(worlds)
    tuned = F.Batch(worlds)
    tuned.set_online_tuning(5)
    ref.step(steps=1)
    tuned.step(steps=1)
    for k in (7, 3, 9):
        ref.step(steps=k)
        tuned.step(steps=k)
    assert ref.sync() == 0 and tuned.sync() == 0
    for c in tuned.size_classes():
        assert c["lanes"] in ((256, 512, 1024) if c["limit"] > 512 else (64, 128, 256)), c
    for w in range(len(worlds)):
        assert ref.export(w) == tuned.export(w), w


def test_step_subset_leaves_other_worlds(gpu):
    worlds = scenes.config3(worlds=6, rows_lo=5, rows_hi=8)
    b = F.Batch(worlds)
    b.step(steps=2)
    assert b.sync() == 0
    before = [b.bodies(w).copy() for w in range(6)]
    b.step_worlds([1, 4], steps=3)
    assert b.sync() == 0
    for w in range(6):
        same = np.array_equal(b.bodies(w), before[w])
        assert same == (w not in (1, 4)), w


def kinetic(bodies, masses):
    v2 = bodies[:, 4] ** 2 + bodies[:, 5] ** 2
    return float(np.sum(0.5 * masses * v2))


def test_config1_pyramid_trajectory_invariants(gpu):
    """820-box pyramid, 256 steps: the top box stays at its height (no collapse),
    no box drops below the ground surface, and kinetic energy decays."""
    worlds = scenes.config1()
    b = F.Batch(worlds)
    init = b.bodies(0).copy()
    ke = []
    for _ in range(8):
        b.step(steps=32)
        assert b.sync() == 0
        cur = b.bodies(0)
        ke.append(kinetic(cur[1:], np.ones(len(cur) - 1)))
    cur = b.bodies(0)
    dyn = cur[1:]
    assert np.all(dyn[:, 1] > 0.4), "a box sank into the ground"
    top0 = init[1:, 1].max()
    # soft contacts compress a 40-row stack by ~3 mm per contact; a collapse drops it by metres
    assert abs(dyn[:, 1].max() - top0) < 0.25, (dyn[:, 1].max(), top0)
    assert np.max(np.abs(dyn[:, 0] - init[1:, 0])) < 0.1
    assert ke[-1] < 1e-2, ke


def test_config3_batch_trajectory_invariants(gpu):
    worlds = scenes.config3(worlds=64)
    b = F.Batch(worlds)
    tops = [b.bodies(w)[1:, 1].max() for w in range(len(worlds))]
    b.step(steps=256)
    assert b.sync() == 0
    for w in range(len(worlds)):
        cur = b.bodies(w)[1:]
        assert np.all(cur[:, 1] > 0.4), w
        assert abs(cur[:, 1].max() - tops[w]) < 0.25, w
        assert np.all(np.abs(cur[:, 4:6]) < 0.5), w


def test_container_no_tunnelling(gpu):
    """Mixed shapes dropped into a segment container stay inside it."""
    worlds = [scenes.container_world(300, 15.0, 30.0, 99, continuous=False, chains=2, drop_band=(0.1, 0.5))]
    b = F.Batch(worlds)
    b.step(steps=240)
    assert b.sync() == 0
    cur = b.bodies(0)[1:]
    assert np.all(cur[:, 1] > -0.5)
    assert np.all(np.abs(cur[:, 0]) < 15.5)
    speed = np.sqrt(cur[:, 4] ** 2 + cur[:, 5] ** 2)
    assert np.median(speed) < 1.0
