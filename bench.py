"""Benchmark: body-substeps/s of the batched MI355X world step, beside the
one-world-at-a-time GPU path and the reference CPU step; p50/p99 step latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config4|config3|config1|config2]

The default workload is the batched heterogeneous scene the throughput target is
stated on: 16,384 worlds of 10-500 mixed bodies with revolute chains. A step is
one b2World_Step of every world (dt = 1/60, 4 substeps): collide, tree, solve
and continuous-collision launches per size class, one workgroup per world, on
inputs already resident in HBM. Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import forge2d_amd as F  # noqa: E402
from forge2d_amd import scenes  # noqa: E402

METRIC = "body-substeps/s: batched vs per-world GPU vs host CPU; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
SUBSTEPS = 4
DT = 1.0 / 60.0
LATENCY_STEPS = 32   # steps timed one at a time for the latency percentiles
EV_STEPS = 32        # steps of the per-kernel HIP-event timing

WORKLOADS = {
    "config3": ("4,096 independent worlds, each a 50x1 ground box plus a pyramid of U{5..30} rows of 0.5 "
                "half-extent boxes; FP32, dt=1/60, 4 substeps, sleep off", lambda n: scenes.config3(worlds=n or 4096)),
    "config4": ("16,384 independent worlds, body counts log-uniform in [10,500]: circles, capsules, 3-8 vertex "
                "polygons and 0-8 revolute chains in a segment container; FP32, dt=1/60, 4 substeps, sleep off",
                lambda n: scenes.config4(worlds=n or 16384)),
    "config1": ("single world: 50x1 ground box and a 40-row pyramid of 820 boxes; FP32, dt=1/60, 4 substeps, "
                "sleep off", lambda n: scenes.config1()),
    "config2": ("single world: 200x200 segment container with 20,000 mixed bodies; FP32, dt=1/60, 4 substeps, "
                "sleep off", lambda n: scenes.config2()),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(st: dict) -> dict:
    """Minimal FP32 field traffic of one step when every stage is a full pass over
    HBM, split by kernel.

    solveKernel, per substep: 88 B/body (integrate velocities 44 + positions 44),
    380 + 128 p per contact constraint (warm start 84 + 24 p, solve and relax
    148 + 52 p each), 576 per revolute joint; per step: 136 B/body and 112 B/shape
    (finalize: body state + transform write, shape AABB + fat AABB), 148 + 100 p per
    constraint (prepare 140 + 72 p, store 8 + 28 p), 148 + 52 p per constraint with
    restitution, 120 per joint (prepare).
    collideKernel, per step: 116 B/shape (fat-AABB sort-and-sweep and move flags),
    276 + 62 p per contact (two transforms, shape geometry, old and new manifold).
    treeKernel, per step: 148 B per proxy of the kinematic and dynamic trees (its
    leaf and one internal node, 32 B each, read and written; the shape's 16-B fat
    box read; its 4-B query rank written).
    The three add up to S.[88 Nb + 380 C + 128 P + 576 Nj] + [136 Nb + 228 Ns + 424 C
    + 162 P + 148 Cr + 52 Pr + 120 Nj + 148 Np]."""
    Nb, Ns, C, P = st["bodies"], st["shapes"], st["contacts"], st["points"]
    Cr, Pr, Nj = st["restitution_contacts"], st["restitution_points"], st["joints"]
    per_sub = 88 * Nb + 380 * C + 128 * P + 576 * Nj
    solve = SUBSTEPS * per_sub + 136 * Nb + 112 * Ns + 148 * C + 100 * P + 148 * Cr + 52 * Pr + 120 * Nj
    collide = 116 * Ns + 276 * C + 62 * P
    tree = 148 * st["tree_proxies"]
    return {"solveKernel": float(solve), "collideKernel": float(collide), "treeKernel": float(tree),
            "step": float(solve + collide + tree)}


def load_pmc(workload: str):
    path = os.path.join(ROOT, "profiles", f"pmc_{workload}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def host_cpu() -> dict:
    """The host's CPU model, logical CPU count and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    # the job's CPU share: OMP_NUM_THREADS where the host sets it, else the affinity set
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": affinity,
            "job_cpu_share": min(share, affinity)}


def _time_reference(harness: str, path: str, steps: int, warmup: int, threads: int, worlds: int) -> dict:
    out = subprocess.run([harness, "time", path, str(steps), str(warmup), str(threads), "0", str(worlds)],
                         capture_output=True, text=True, timeout=900, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(spec: bytes, worlds: int, budget_s: float, warmup: int, steps: int):
    """Time the reference Box2D (built from its sources by oracle/Makefile, hook
    flags) on this host on a bounded sample of the same worlds, over the same step
    window as the GPU's timed region: each world is stepped `warmup` steps untimed,
    then `steps` steps timed. Once on one thread, once on the job's CPU share
    (threads stepping disjoint worlds). The sample is the first worlds of the
    workload (the generator draws sizes at random), sized so that each run takes
    about budget_s seconds."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        return None
    host = host_cpu()
    share = max(1, host["job_cpu_share"])
    runs = {}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "scene.bin")
        with open(path, "wb") as f:
            f.write(spec)
        for threads in sorted({1, share}):
            # calibrate on a small sample, then size the sample to the time budget
            n0 = min(worlds, 2 * threads)
            r = _time_reference(harness, path, steps, warmup, threads, n0)
            per_world = r["wall_s"] / n0 * threads
            n = min(worlds, max(n0, int(budget_s * threads / max(per_world, 1e-6))))
            if n > n0:
                r = _time_reference(harness, path, steps, warmup, threads, n)
            runs[threads] = r
    t1, tn = runs[1], runs[share]
    used = int(tn["threads"])
    window = f"steps {warmup}..{warmup + steps - 1} of each world (the GPU's timed window), {warmup} untimed steps first"
    return {"value": tn["body_substeps_per_s"], "unit": "body-substeps/s", "cores": used, "kind": "reference",
            "sample": f"first {tn['worlds']} worlds of the workload ({int(tn['bodies'])} bodies), {window}, "
                      f"{used} threads stepping disjoint worlds; reference Box2D v3.1.1 built with the forge2d hook "
                      f"flags (-std=c17 -O2, SSE2)",
            "step_ms_p50": tn["step_ms_p50"], "step_ms_p99": tn["step_ms_p99"], "wall_s": tn["wall_s"],
            "t1": {"value": t1["body_substeps_per_s"], "cores": 1, "worlds": t1["worlds"],
                   "bodies": int(t1["bodies"]), "wall_s": t1["wall_s"], "step_ms_p50": t1["step_ms_p50"]},
            "host": host,
            "all_host_cpus_linear_bound": t1["body_substeps_per_s"] * (host["os_cpu_count"] or 1),
            "note": f"threads limited to the job's CPU share ({share}); all_host_cpus_linear_bound = the one-thread "
                    f"rate x os.cpu_count(), an upper bound for every host CPU stepping disjoint worlds"}


def ray_queries(spec: bytes, rays_per_world: int, seed: int = 17) -> np.ndarray:
    """Closest-ray queries (f2dQuery records), rays_per_world per world in world order:
    origins uniform over the box of the world's initial body positions (1 m margin),
    directions uniform, lengths uniform in [0.1, 1] x the box diagonal."""
    parsed = scenes.parse(spec)
    lo = np.array([[b["px"].min(), b["py"].min()] for _, b, _, _ in parsed], np.float64) - 1.0
    hi = np.array([[b["px"].max(), b["py"].max()] for _, b, _, _ in parsed], np.float64) + 1.0
    rng = np.random.default_rng(seed)
    world = np.repeat(np.arange(len(parsed)), rays_per_world)
    n = len(world)
    origin = lo[world] + rng.random((n, 2)) * (hi - lo)[world]
    angle = rng.uniform(-np.pi, np.pi, n)
    length = rng.uniform(0.1, 1.0, n) * np.linalg.norm(hi - lo, axis=1)[world]
    q = np.zeros(n, F.QUERY_DTYPE)
    q["world"], q["kind"], q["cat_lo"], q["mask_lo"], q["mask_hi"] = world, F.Q_RAY_CLOSEST, 1, F.ALL_BITS, F.ALL_BITS
    q["p"][:, 0], q["p"][:, 1] = origin[:, 0], origin[:, 1]
    q["p"][:, 2], q["p"][:, 3] = length * np.cos(angle), length * np.sin(angle)
    return q


def measure_queries(args, batch, spec: bytes, W: int):
    """Closest-ray throughput of the query kernel on the stepped batch (queries, hits and
    counts resident in HBM, one launch per pass), beside the reference's
    b2World_CastRayClosest on a sample of the same worlds and rays."""
    import torch

    q = ray_queries(spec, args.query_rays)
    n = len(q)
    dq = torch.from_numpy(q.view(np.uint8).copy()).cuda()
    dh = torch.zeros(n * F.HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    dc = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    L = F.lib()

    def launch():
        if L.f2dg_query_device(batch.h, dq.data_ptr(), n, dh.data_ptr(), dc.data_ptr(), 1) != 0:
            raise RuntimeError("query launch failed")

    launch()
    batch.synchronize()
    reps = 10
    a = time.perf_counter()
    for _ in range(reps):
        launch()
    batch.synchronize()
    ms = 1e3 * (time.perf_counter() - a) / reps
    counts = dc.cpu().numpy()
    out = {"kind": "closest ray (b2World_CastRayClosest)", "rays": n, "rays_per_world": args.query_rays,
           "ms_per_launch": ms, "rays_per_s": n / (ms * 1e-3), "hit_rate": float((counts > 0).mean()),
           "kernel": "queryKernel",
           "sample": f"{args.query_rays} seeded rays per world over the world's box, after the bench's steps; one "
                     f"lane per ray walking the world's three trees (pending enlargements applied and rays grouped "
                     f"by world in the same launch sequence)"}
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not args.no_cpu and os.path.exists(harness):
        m = min(W, 1024)
        threads = max(1, host_cpu()["job_cpu_share"])
        with tempfile.TemporaryDirectory() as td:
            sp, qp = os.path.join(td, "scene.bin"), os.path.join(td, "q.bin")
            with open(sp, "wb") as f:
                f.write(spec)
            with open(qp, "wb") as f:
                f.write(q[q["world"] < m].tobytes())
            r = subprocess.run([harness, "qtime", sp, qp, "60", str(threads), "0", str(m)], capture_output=True,
                               text=True, timeout=600, check=True)
            c = json.loads(r.stdout.strip().splitlines()[-1])
        out["cpu_baseline"] = {"value": c["rays_per_s"], "unit": "rays/s", "cores": c["threads"], "kind": "reference",
                               "sample": f"first {m} worlds, the same {c['rays']} rays after 60 steps, "
                                         f"{c['threads']} threads querying disjoint worlds, reference Box2D v3.1.1 "
                                         f"dynamic-tree ray casts (-std=c17 -O2)", "hit_rate": c["hits"] / max(1, c["rays"])}
        out["vs_cpu"] = out["rays_per_s"] / c["rays_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="config4", choices=sorted(WORKLOADS))
    ap.add_argument("--worlds", type=int, default=0, help="override the world count (testing)")
    ap.add_argument("--per-world-sample", type=int, default=64)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="run the serving scenario (Poisson arrivals, continuous batching) instead")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-worlds", type=int, default=2000)
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    ap.add_argument("--serve-max-backlog", type=float, default=2.0, metavar="SECONDS",
                    help="admission control: refuse requests that would queue more than this many seconds "
                         "of work at the measured throughput (0: admit everything)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain elementwise copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--kernel-profile", action="store_true",
                    help="run only the per-kernel timing phase (size classes launched one after another, "
                         "W + K steps) so a kernel trace of this command holds exactly those dispatches")
    ap.add_argument("--lanes", default="",
                    help="comma list of workgroup lanes per size class instead of the autotuner's choice")
    ap.add_argument("--phase-times", type=int, default=0, metavar="STEPS",
                    help="after the timed steps, time the step's phases as separate launches over STEPS steps")
    ap.add_argument("--query-rays", type=int, default=64, metavar="N",
                    help="closest-ray queries per world timed after the step measurement (0: none)")
    ap.add_argument("--solve-lds", default="",
                    help="comma list of solve-launch LDS bytes per size class (with --lanes)")
    ap.add_argument("--tune-period", type=int, default=64, metavar="STEPS",
                    help="online lane tuning: a three-step trial round every STEPS steps (0: off)")
    ap.add_argument("--full-steps", type=int, default=512, metavar="N",
                    help="also step a fresh batch of the workload N steps from its initial state (0: skip)")
    ap.add_argument("--trace-summary", default="", metavar="CSV",
                    help="summarise a rocprofv3 kernel trace of a bench run over its kernel-timing window and exit")
    ap.add_argument("--trace-dispatches", type=int, default=5, metavar="N",
                    help="dispatches per step kernel per step in the traced run (its size-class count)")
    ap.add_argument("--pmc-summary", nargs=2, default=None, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/pmc_<workload>.json from two counter runs of --kernel-profile and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        mib = args.copy_calibration or 512
        cmd = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE -- python3 -u bench.py --kernel-profile --workload "
               f"{args.workload} --copy-calibration {mib}" + (f" --lanes {args.lanes}" if args.lanes else "") +
               (f" --solve-lds {args.solve_lds}" if args.solve_lds else ""))
        res = pmc_summary(args.pmc_summary[0], args.pmc_summary[1], EV_STEPS, args.trace_dispatches,
                          args.workload, cmd, mib)
        with open(os.path.join(ROOT, "profiles", f"pmc_{args.workload}.json"), "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res, indent=1))
        return
    if args.trace_summary:
        trace_summary(args.trace_summary, EV_STEPS, args.trace_dispatches)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: every world of a batch lives on a single device and worlds never "
              f"exchange data; --gpus {args.gpus} is not supported (shard worlds across processes instead)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.serve > 0:
        serve(args)
        return

    if args.query_rays > 0 and not args.kernel_profile:
        # the query measurement keeps its rays and hits in torch tensors: torch's HIP
        # context comes up before the library's first device call
        import torch

        torch.zeros(1, device="cuda")
    desc, make = WORKLOADS[args.workload]
    t0 = time.time()
    worlds = make(args.worlds)
    spec = scenes.to_bytes(worlds)
    del worlds
    log(f"[bench] {args.workload}: spec {len(spec) / 1e6:.1f} MB built in {time.time() - t0:.1f}s")

    batch = F.Batch(spec)
    W = batch.world_count
    awake = batch.awake_bodies()
    batch.upload()
    log(f"[bench] uploaded {W} worlds, {awake} awake bodies, {batch.device_bytes() / 2**30:.2f} GiB device")

    if args.lanes:
        for i, lanes in enumerate(int(x) for x in args.lanes.split(",")):
            batch.set_class_block(i, lanes)
    if args.solve_lds:
        for i, nbytes in enumerate(int(x) for x in args.solve_lds.split(",")):
            batch.set_class_solve_lds(i, nbytes)
    if args.kernel_profile:
        kernel_profile(args, batch, awake, desc)
        return

    batch.step(DT, SUBSTEPS, steps=args.warmup)
    batch.synchronize()
    if batch.error_bits() != 0:
        raise RuntimeError("device reported a capacity error after warmup")
    st = batch.stats()
    log(f"[bench] after warmup: {st}")
    tune_ms = None
    tune_steps = 0
    if not args.lanes:
        _, tune_ms = batch.autotune(DT, SUBSTEPS, steps=4)
        tune_steps = 9 * 4
        batch.synchronize()
        log(f"[bench] autotuned workgroup lanes per size class: "
            f"{[c['lanes'] for c in batch.size_classes()]} (whole-batch ms per 4 steps at one size: {tune_ms})")
    if args.tune_period > 0 and not args.lanes:
        batch.set_online_tuning(args.tune_period)
    start = args.warmup + tune_steps

    # timed region: K steps, synchronized on both sides
    batch.synchronize()
    t1 = time.perf_counter()
    batch.step(DT, SUBSTEPS, steps=args.steps)
    batch.synchronize()
    wall = time.perf_counter() - t1
    ms_per_step = 1e3 * wall / args.steps
    value = awake * SUBSTEPS * args.steps / wall
    st = batch.stats()
    log(f"[bench] batched: {ms_per_step:.3f} ms/step, {value:.4e} body-substeps/s (steps {start}..{start + args.steps - 1})")

    phase_ms = None
    if args.phase_times > 0:
        phase_ms = {k: v / args.phase_times for k, v in batch.phase_times(DT, SUBSTEPS, args.phase_times).items()}
        log(f"[bench] phase ms per step: {phase_ms}")

    # per-step latency percentiles (each step synchronized on its own)
    lat = []
    for _ in range(LATENCY_STEPS):
        a = time.perf_counter()
        batch.step(DT, SUBSTEPS, steps=1)
        batch.synchronize()
        lat.append(1e3 * (time.perf_counter() - a))
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    if batch.error_bits() != 0:
        raise RuntimeError("device reported a capacity error")

    # one-world-at-a-time GPU path on a sample of the same worlds
    per_world = None
    if W > 1 and args.per_world_sample > 0:
        n = min(W, args.per_world_sample)
        sub = F.Batch(spec_subset(spec, n))
        sub_awake = sub.awake_bodies()
        sub.upload()
        sub.step_each(DT, SUBSTEPS, steps=1)
        sub.synchronize()
        k = 4
        a = time.perf_counter()
        sub.step_each(DT, SUBSTEPS, steps=k)
        sub.synchronize()
        t = time.perf_counter() - a
        per_world = {"value": sub_awake * SUBSTEPS * k / t, "unit": "body-substeps/s",
                     "sample": f"first {n} worlds, {k} steps each, one launch per world per step",
                     "ms_per_world_step": 1e3 * t / (n * k)}
        sub.close()
        log(f"[bench] per-world GPU path: {per_world}")

    # the workload's whole run: a fresh batch stepped from the initial state
    full = full_run(spec, args.full_steps, args.tune_period if not args.lanes else 0) if args.full_steps > 0 else None

    # kernel durations with HIP events on the launch stream (size classes one after
    # another): the last step-kernel dispatches of this process, so a kernel trace of
    # this command holds this window as its last EV_STEPS x classes dispatches per kernel
    ev_first = start + args.steps + LATENCY_STEPS
    launch_ms = {k: v / EV_STEPS for k, v in batch.kernel_times(DT, SUBSTEPS, steps=EV_STEPS).items()}
    if batch.error_bits() != 0:
        raise RuntimeError("device reported a capacity error")
    st_ev = batch.stats()
    nbytes = algorithmic_bytes(st_ev)
    dominant = max(launch_ms, key=launch_ms.get)
    achieved = nbytes[dominant] / (launch_ms[dominant] * 1e-3) / 1e9
    step_gbs = nbytes["step"] / (sum(launch_ms.values()) * 1e-3) / 1e9
    pmc = load_pmc(args.workload)
    traffic = pmc.get(dominant, {}).get("hbm_bytes_per_launch") if pmc else None
    log(f"[bench] kernel ms per step (steps {ev_first}..{ev_first + EV_STEPS - 1}) {launch_ms}, "
        f"algorithmic bytes {nbytes}, {dominant} {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(spec, W, args.cpu_budget, start, min(args.steps, 64))
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    queries = None
    if args.query_rays > 0:
        queries = measure_queries(args, batch, spec, W)
        log(f"[bench] queries: {queries}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded scene generator (forge2d_amd/scenes.py), no sleep",
        "config": {"workload": f"{args.workload}: {desc}", "worlds": W, "awake_bodies": awake,
                   "substeps": SUBSTEPS, "dt": DT, "timed_steps": [start, start + args.steps - 1],
                   "contacts": st["contacts"], "contacts_per_body": st["contacts"] / max(awake, 1),
                   "manifold_points": st["points"], "joints": st["joints"], "size_classes": batch.size_classes(),
                   "autotune_ms_per_4_steps": tune_ms, "autotune_steps": tune_steps,
                   "online_tuning_period": args.tune_period if not args.lanes else 0,
                   "lanes_override": args.lanes or None,
                   "parallelism": "one workgroup per world, worlds launched in size classes (own LDS size and workgroup "
                                  "size, own stream), largest first; collide, tree, solve and continuous-collision "
                                  "launches per class per step"},
        "latency_ms": {"p50": p50, "p99": p99, "samples": len(lat)},
        "phase_ms_per_step": phase_ms,
        "per_world_gpu": per_world,
        "full_run": full,
        "batched_vs_per_world": (value / per_world["value"]) if per_world else None,
        "batched_vs_cpu": (value / cpu["value"]) if cpu else None,
        "batched_vs_cpu_one_thread": (value / cpu["t1"]["value"]) if cpu else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dominant,
                     "launch_ms": launch_ms[dominant], "algorithmic_bytes_per_launch": nbytes[dominant],
                     "window": {"steps": [ev_first, ev_first + EV_STEPS - 1],
                                "dispatches_per_step": len(batch.size_classes()),
                                "contacts_per_body": st_ev["contacts"] / max(awake, 1),
                                "trace": "the last EV_STEPS x dispatches_per_step dispatches of each step kernel "
                                         "in a kernel trace of this command"},
                     "other_kernel": {k: {"launch_ms": launch_ms[k], "algorithmic_bytes_per_launch": nbytes.get(k)}
                                      for k in launch_ms if k != dominant},
                     "step_achieved_gbs": step_gbs},
        "cpu_baseline": cpu,
        "queries": queries,
    }
    print(json.dumps(out), flush=True)


def kernel_profile(args, batch, awake, desc):
    """The step kernels alone, for counter runs: the same step window as the default
    run's kernel timing (warm-up, the autotuner's 36 trial steps, the timed steps and
    the latency steps, stepped untimed), then EV_STEPS steps with the size classes
    launched one after another on one stream, each kernel bracketed by HIP events.
    Those are the last step-kernel dispatches of the run: EV_STEPS x classes per
    kernel."""
    pre = args.warmup + 9 * 4 + args.steps + LATENCY_STEPS
    batch.step(DT, SUBSTEPS, steps=pre)
    batch.synchronize()
    if batch.error_bits() != 0:
        raise RuntimeError("device reported a capacity error")
    launch_ms = {k: v / EV_STEPS for k, v in batch.kernel_times(DT, SUBSTEPS, steps=EV_STEPS).items()}
    if batch.error_bits() != 0:
        raise RuntimeError("device reported a capacity error")
    st = batch.stats()
    nbytes = algorithmic_bytes(st)
    dominant = max(launch_ms, key=launch_ms.get)
    achieved = nbytes[dominant] / (launch_ms[dominant] * 1e-3) / 1e9
    classes = batch.size_classes()
    pmc = load_pmc(args.workload)
    out = {
        "metric": METRIC, "mode": "kernel-profile", "n_gpus": 1, "steps": EV_STEPS, "warmup": pre,
        "dtype": "f32", "config": {"workload": f"{args.workload}: {desc}", "awake_bodies": awake,
                                   "size_classes": classes, "stats": st},
        "kernel_ms_per_step": launch_ms,
        "dispatches_per_kernel": {k: len(classes) * EV_STEPS for k in launch_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc.get(dominant, {}).get("hbm_bytes_per_launch") if pmc else None,
                     "kernel": dominant, "launch_ms": launch_ms[dominant],
                     "algorithmic_bytes_per_launch": nbytes[dominant], "window": [pre, pre + EV_STEPS - 1]},
    }
    log(f"[bench] kernel profile: {launch_ms} ms per step, {dominant} {achieved:.0f} GB/s")
    print(json.dumps(out), flush=True)


def full_run(spec: bytes, steps: int, tune_period: int) -> dict:
    """The workload's whole run on a fresh batch from its initial state: `steps`
    steps in chunks of 64, each chunk synchronized and timed, with the touching
    contacts per body after the first step and at the end (online lane tuning as in
    the timed region, no up-front autotuning)."""
    b = F.Batch(spec)
    awake = b.awake_bodies()
    b.upload()
    if tune_period > 0:
        b.set_online_tuning(tune_period)
    b.step(DT, SUBSTEPS, steps=1)
    b.synchronize()
    c0 = b.stats()["contacts"]
    chunks = []
    done = 1
    total_ms = 0.0
    while done < steps:
        k = min(64, steps - done)
        a = time.perf_counter()
        b.step(DT, SUBSTEPS, steps=k)
        b.synchronize()
        ms = 1e3 * (time.perf_counter() - a)
        chunks.append(ms / k)
        total_ms += ms
        done += k
    if b.error_bits() != 0:
        raise RuntimeError("device reported a capacity error in the full run")
    c1 = b.stats()["contacts"]
    b.close()
    out = {"steps": steps, "timed_steps": [1, steps - 1], "ms_per_step": total_ms / (steps - 1),
           "value": awake * SUBSTEPS * (steps - 1) / (total_ms * 1e-3), "unit": "body-substeps/s",
           "ms_per_step_per_64_steps": chunks, "contacts_per_body_step1": c0 / max(awake, 1),
           "contacts_per_body_end": c1 / max(awake, 1)}
    log(f"[bench] full run: {out}")
    return out


def pmc_summary(fetch_csv: str, write_csv: str, window_steps: int, dispatches_per_step: int, workload: str,
                command: str, mib: int) -> dict:
    """HBM bytes per step of each step kernel from two rocprofv3 counter runs of
    `bench.py --kernel-profile --copy-calibration MIB` (FETCH_SIZE, then WRITE_SIZE;
    KiB per dispatch): each kernel's last window_steps x dispatches_per_step
    dispatches summed and divided by window_steps, each counter scaled by what it
    reads for the plain copy of known size that the run starts with."""
    import csv

    def per_dispatch(path, counter):
        vals = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                key = int(r["Dispatch_Id"])
                name = r.get("Kernel_Name", "")
                old = vals.get(key, (name, 0.0))
                vals[key] = (name, old[1] + float(r["Counter_Value"]))
        return [vals[k] for k in sorted(vals)]

    out = {"workload": workload, "command": command, "window_steps": window_steps,
           "dispatches_per_step": dispatches_per_step}
    scales = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = per_dispatch(path, counter)
        copies = [v for n, v in rows if "elementwise" in n and "Mul" in n]
        # the copy reads and writes mib MiB: the counter's reading per KiB moved
        scales[counter] = (mib * 1024.0) / (sum(copies) / len(copies)) if copies else None
        fams = {}
        for n, v in rows:
            base = n.split("(")[0].replace("void ", "").split("<")[0].strip()
            if base.startswith("f2d::"):
                fams.setdefault(base.replace("f2d::", ""), []).append(v)
        for base, vs in fams.items():
            last = vs[-window_steps * dispatches_per_step:]
            key = "fetch_kib_per_step" if counter == "FETCH_SIZE" else "write_kib_per_step"
            out.setdefault(base, {})[key] = sum(last) / window_steps
    out["fetch_scale"], out["write_scale"] = scales["FETCH_SIZE"], scales["WRITE_SIZE"]
    for base, v in out.items():
        if isinstance(v, dict) and "fetch_kib_per_step" in v and "write_kib_per_step" in v and None not in scales.values():
            v["hbm_bytes_per_launch"] = 1024.0 * (v["fetch_kib_per_step"] * scales["FETCH_SIZE"] +
                                                  v["write_kib_per_step"] * scales["WRITE_SIZE"])
    return out


def trace_summary(path: str, window_steps: int, dispatches_per_step: int):
    """Per-step kernel time from a rocprofv3 kernel trace (--kernel-trace -f csv) of a
    bench command: per step kernel (the name without its template arguments), the
    summed duration of its last window_steps x dispatches_per_step dispatches, divided
    by window_steps; with every dispatch's average beside it."""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("KernelName") or ""
            t0 = int(r.get("Start_Timestamp") or r.get("BeginNs"))
            t1 = int(r.get("End_Timestamp") or r.get("EndNs"))
            rows.append((t0, name, t1 - t0))
    rows.sort()
    fams = {}
    for t0, name, d in rows:
        base = name.split("(")[0].replace("void ", "").split("<")[0].strip()
        fams.setdefault(base, []).append(d)
    n = window_steps * dispatches_per_step
    print("| kernel | dispatches in the run | avg ns per dispatch | window dispatches | window ms per step |")
    print("|---|---|---|---|---|")
    for base, ds in sorted(fams.items(), key=lambda kv: -sum(kv[1])):
        if not base.startswith("f2d::"):
            continue
        last = ds[-n:] if len(ds) >= n else ds
        print(f"| {base} | {len(ds)} | {sum(ds) / len(ds):.0f} | {len(last)} | {sum(last) / window_steps / 1e6:.4f} |")


def serve(args):
    """Serving scenario: Poisson arrivals of step requests (k in {1, 8, 60}) over a
    population of worlds log-uniform in [50, 20000] bodies, continuous batching on
    one GPU within an HBM budget. Prints one JSON line."""
    from forge2d_amd import serving

    t0 = time.time()
    pop = serving.Population(args.serve_worlds)
    reqs = serving.poisson_requests(args.serve_rate, args.serve, pop)
    for t in sorted(set(pop.template_of.tolist())):
        pop.spec(int(np.nonzero(pop.template_of == t)[0][0]))
    log(f"[serve] {len(reqs)} requests over {args.serve:.0f}s, population {len(pop.sizes)} worlds, templates "
        f"built in {time.time() - t0:.1f}s")
    eng = serving.ServingEngine(pop, int(args.serve_budget_gib * 2**30),
                                max_backlog_s=args.serve_max_backlog if args.serve_max_backlog > 0 else None)
    t1 = time.time()
    preloaded = eng.preload()
    log(f"[serve] population resident before the first request: {preloaded} ({time.time() - t1:.1f}s)")
    st = eng.run(reqs, horizon=args.serve, log=log)
    st["preloaded"] = preloaded
    eng.close()
    out = {
        "metric": METRIC, "value": st["body_substeps_per_s"], "unit": "body-substeps/s", "n_gpus": 1,
        "steps": st["iterations"], "warmup": 0, "ms_per_step": 1e3 * st["wall_s"] / max(st["iterations"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded container worlds (circles, capsules, polygons), Poisson arrivals",
        "config": {"workload": f"config5 serving: Poisson arrivals at {args.serve_rate:g} req/s for {args.serve:g}s, "
                               f"{args.serve_worlds} worlds log-uniform in [50, 20000] bodies, k in {{1, 8, 60}} "
                               f"steps per request, HBM budget {args.serve_budget_gib:g} GiB, admission control "
                               f"at {args.serve_max_backlog:g} s of queued work",
                   "parallelism": "continuous batching: one step launch set per resident batch per iteration, "
                                  "batches on separate HIP streams"},
        "latency_ms": st["latency_ms"], "serving": st,
    }
    print(json.dumps(out), flush=True)


def copy_calibration(mib: int):
    """A plain copy kernel of known size (reads and writes mib MiB once per launch):
    the ratio of its FETCH_SIZE / WRITE_SIZE counter readings to mib MiB scales the
    same counters read for the physics kernels."""
    import torch
    n = mib * 2**20 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 3 launches of an elementwise copy of {mib} MiB")
    del src, dst


def spec_subset(spec: bytes, n: int) -> bytes:
    worlds = scenes.parse(spec)[:n]
    parts = [scenes.MAGIC, np.array([n, 0], "<i4").tobytes()]
    for h, b, s, j in worlds:
        parts += [h.tobytes(), b.tobytes(), s.tobytes(), j.tobytes()]
    return b"".join(parts)


if __name__ == "__main__":
    main()
