"""Serving loop: continuous batching of world-step requests on one GPU.

Requests arrive over time; each names a world of a population and a number of
steps k. A world that is not resident is created on the device (admission is
bounded by an HBM budget; when the budget is full, idle worlds are evicted in
least-recently-used order, a whole device batch at a time, each world's state
exported to host memory first); a resident world is continued from its device
state, and a re-admitted world from its exported state. A request for a world
that could never fit the budget is refused, and so is one whose world would then
owe more steps than `max_backlog_s` seconds of iterations at the iteration time
measured so far (admission control: latency stays bounded at the rate the engine
admits). `preload` makes the whole population resident before the first
request when it fits the budget (HBM is large enough for most populations). Every scheduler iteration launches one step for
every world that still owes steps — across all resident batches, each on its
own HIP stream so their kernels overlap — and requests that arrive meanwhile
join at the next iteration. A request completes when its k-th step has
finished; its latency is completion time minus arrival time.

Worlds never interact, so a world's trajectory is the same whether it is stepped
alone or in the mix (`tests/test_gpu_serving.py` checks this bit for bit).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, scenes, snapshot


@dataclass
class Request:
    rid: int
    world: int          # population index
    steps: int
    arrival: float      # seconds since the start of the run
    done: float = -1.0
    steps_left: int = -1


@dataclass
class _Resident:
    batch: "Generation"
    index: int          # world index inside its batch
    owed: int = 0       # steps still to run
    last_used: float = 0.0
    steps_done: int = 0  # steps run since the world was created on the device
    queue: collections.deque = field(default_factory=collections.deque)  # requests in order


class Generation:
    """One device batch holding the worlds admitted together."""

    def __init__(self, members, specs, stream=None, states=None):
        self.members = list(members)                      # population indices
        self.batch = Batch(b"".join([scenes.MAGIC, np.array([len(specs), 0], "<i4").tobytes()] + specs))
        # worlds evicted earlier continue from their exported states
        for i, snap in enumerate(states or []):
            if snap is not None:
                self.batch.import_state(i, snap)
        self.stream = stream
        if stream is not None:
            self.batch.set_stream(stream.cuda_stream)
        self.batch.upload()
        self.bytes = self.batch.device_bytes()
        self.bodies = [self.batch.body_count(i) for i in range(len(specs))]
        self.alive = len(specs)

    def close(self):
        self.batch.close()


def world_spec(w: scenes.World) -> bytes:
    """The per-world part of a scene spec (header, bodies, shapes, joints)."""
    full = scenes.to_bytes([w])
    return full[len(scenes.MAGIC) + 8:]


class Population:
    """World sizes log-uniform in [lo, hi] bodies; worlds are the container mix
    (circles, capsules, polygons) built from a few hundred seeded templates."""

    def __init__(self, count: int, lo: int = 50, hi: int = 20000, templates: int = 96, seed: int = 5):
        rng = np.random.default_rng(seed)
        self.sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).round().astype(int)
        edges = np.exp(np.linspace(math.log(lo), math.log(hi), templates))
        self.template_of = np.clip(np.searchsorted(edges, self.sizes), 0, templates - 1)
        self._edges = edges
        self._cache: dict[int, bytes] = {}

    def spec(self, world: int) -> bytes:
        t = int(self.template_of[world])
        if t not in self._cache:
            n = int(round(self._edges[t]))
            side = max(8.0, math.sqrt(n) * 1.2)
            self._cache[t] = world_spec(scenes.container_world(n, side, 2.0 * side, scenes.stream_seed(5, t),
                                                               continuous=True, drop_band=(0.05, 0.95)))
        return self._cache[t]


class ServingEngine:
    def __init__(self, population: Population, hbm_budget: int, dt: float = 1.0 / 60.0, substeps: int = 4,
                 use_streams: bool = True, max_backlog_s: float | None = None, iteration_s: float = 0.02):
        self.pop = population
        self.max_backlog_s = max_backlog_s
        self.iteration_s = iteration_s  # mean wall time of a scheduler iteration, measured as it runs
        self.budget = hbm_budget
        self.dt, self.substeps = dt, substeps
        self.resident: dict[int, _Resident] = {}
        self.generations: list[Generation] = []
        self.used = 0
        self.use_streams = use_streams
        self.body_substeps = 0
        self.iterations = 0
        self.evicted = 0
        self.refused = 0
        self.saved: dict[int, bytes] = {}        # exported states of evicted worlds
        self.total_steps: dict[int, int] = collections.defaultdict(int)  # steps run per world, all admissions

    # -- admission -------------------------------------------------------------------------------------------
    def _estimate(self, world: int) -> int:
        # conservative device bytes of a world: ~3.3 KB per body measured (contact
        # slots, constraint lanes and pair scratch at 6 per shape dominate)
        return 65536 + 4096 * int(self.pop.sizes[world])

    def _evict_until(self, need: int, now: float) -> bool:
        idle = [g for g in self.generations if all(self.resident[m].owed == 0 for m in g.members
                                                    if m in self.resident)]
        idle.sort(key=lambda g: max(self.resident[m].last_used for m in g.members if m in self.resident))
        for g in idle:
            if self.used + need <= self.budget:
                break
            if any(m in self.resident and self.resident[m].queue for m in g.members):
                continue
            for m in g.members:
                res = self.resident.pop(m, None)
                if res is not None:
                    self.saved[m] = g.batch.export(res.index)
            self.used -= g.bytes
            self.generations.remove(g)
            g.close()
            self.evicted += 1
        return self.used + need <= self.budget

    def _admit(self, worlds: list[int], now: float):
        need = sum(self._estimate(w) for w in worlds)
        if self.used + need > self.budget and not self._evict_until(need, now):
            return False
        stream = None
        if self.use_streams:
            import torch

            stream = torch.cuda.Stream()
        g = Generation(worlds, [self.pop.spec(w) for w in worlds], stream, [self.saved.pop(w, None) for w in worlds])
        self.generations.append(g)
        self.used += g.bytes
        for i, w in enumerate(worlds):
            self.resident[w] = _Resident(g, i, 0, now)
        return True

    # -- scheduling ------------------------------------------------------------------------------------------
    def run(self, requests: list[Request], horizon: float, log=None) -> dict:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: collections.deque = collections.deque()
        finished: list[Request] = []
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                if self._estimate(r.world) > self.budget:
                    self.refused += 1  # the world can never be resident
                    continue
                if self.max_backlog_s is not None:
                    res = self.resident.get(r.world)
                    owed = (res.owed if res is not None else 0) + sum(q.steps for q in waiting if q.world == r.world)
                    if (owed + r.steps) * self.iteration_s > self.max_backlog_s:
                        self.refused += 1  # over capacity: refused rather than queued
                        continue
                waiting.append(r)
            # admit new worlds (grouped into one batch per iteration), queue steps on resident ones
            new_worlds = []
            still = collections.deque()
            for r in waiting:
                if r.world in self.resident or r.world in new_worlds:
                    continue
                new_worlds.append(r.world)
            if new_worlds and not self._admit(new_worlds, now):
                # budget full of busy worlds: admit what fits one by one, refuse the rest for now
                for w in list(new_worlds):
                    if not self._admit([w], now):
                        break
            for r in waiting:
                res = self.resident.get(r.world)
                if res is None:
                    still.append(r)
                    continue
                res.queue.append(r)
                res.owed += r.steps
            waiting = still
            active = [(w, res) for w, res in self.resident.items() if res.owed > 0]
            if not active:
                if not pending and not waiting:
                    break
                if waiting:
                    # nothing runs and the waiting worlds could not be admitted even
                    # with every idle batch evicted: refuse them rather than spin
                    self.refused += len(waiting)
                    waiting = collections.deque()
                    continue
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.001)))
                continue
            # one step for every world that owes steps: one launch set per batch, batches overlap
            per_gen: dict[int, list] = collections.defaultdict(list)
            for w, res in active:
                per_gen[id(res.batch)].append((w, res))
            launched = []
            for items in per_gen.values():
                g = items[0][1].batch
                g.batch.step_worlds([res.index for _, res in items], self.dt, self.substeps, 1)
                launched.append(g)
            for g in launched:
                g.batch.synchronize()
            t_prev = now
            now = time.perf_counter() - t0
            self.iterations += 1
            self.iteration_s = 0.9 * self.iteration_s + 0.1 * (now - t_prev)
            for w, res in active:
                res.owed -= 1
                res.steps_done += 1
                self.total_steps[w] += 1
                res.last_used = now
                self.body_substeps += res.batch.bodies[res.index] * self.substeps
                # the head request completes when its own steps are done
                if res.queue:
                    head = res.queue[0]
                    if head.steps_left < 0:
                        head.steps_left = head.steps
                    head.steps_left -= 1
                    if head.steps_left == 0:
                        head.done = now
                        finished.append(res.queue.popleft())
            if log and self.iterations % 200 == 0:
                log(f"[serve] t={now:.1f}s iter={self.iterations} resident={len(self.resident)} "
                    f"active={len(active)} done={len(finished)} used={self.used / 2**30:.2f} GiB")
            if now > horizon and not pending:
                # drain: stop admitting, finish what is queued
                if not waiting and not any(res.owed for res in self.resident.values()):
                    break
        wall = time.perf_counter() - t0
        lat = np.array([r.done - r.arrival for r in finished]) * 1e3
        return {
            "requests": len(requests), "completed": len(finished), "wall_s": wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)) if lat.size else None,
                           "p99": float(np.percentile(lat, 99)) if lat.size else None,
                           "mean": float(lat.mean()) if lat.size else None},
            "body_substeps_per_s": self.body_substeps / wall, "iterations": self.iterations,
            "resident_worlds": len(self.resident), "batches": len(self.generations),
            "hbm_used_gib": self.used / 2**30, "evicted_batches": self.evicted, "refused": self.refused,
            "admitted_per_s": (len(requests) - self.refused) / max(horizon, 1e-9),
            "iteration_ms": 1e3 * self.iteration_s,
        }

    def preload(self) -> bool:
        """Make every world of the population resident (one device batch) when the
        whole population fits the budget; returns whether it did."""
        worlds = list(range(len(self.pop.sizes)))
        if sum(self._estimate(w) for w in worlds) > self.budget:
            return False
        return self._admit(worlds, 0.0)

    def world_step(self, world: int) -> int:
        """The world's step counter: from its device batch when resident, else from
        its exported state (0 for a world never admitted)."""
        res = self.resident.get(world)
        if res is not None:
            return res.batch.batch.counters(res.index)["step"]
        snap = self.saved.get(world)
        return snapshot.parse_one(snap).step if snap is not None else 0

    def close(self):
        for g in self.generations:
            g.close()
        self.generations.clear()
        self.resident.clear()


def poisson_requests(rate: float, duration: float, population: Population, seed: int = 11,
                     steps_choices=(1, 8, 60)) -> list[Request]:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        out.append(Request(len(out), int(rng.integers(len(population.sizes))), int(rng.choice(steps_choices)), t))
    return out
