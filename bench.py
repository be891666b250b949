#!/usr/bin/env python3
"""Contact constraint-iterations/s of the soft-step solve path on one MI355X.

Workload: the 32,768-body drop-and-settle pile (half boxes, half spheres, inside four static walls
on a half-space; twofold_amd.scenes.pile). It is stepped `--settle` steps with the host narrowphase
(twofold_amd/csrc/scene.cpp) and the GPU solver; then every further step collides the pile on the
host and uploads its contacts (untimed), and the timed part of the step is one full solve block +
sleep pass over them on the GPU (src/world.cpp:1591-1691, 1174-1247) with the inputs resident, fast
pairs collided again mid-step by the native host callback as the reference's solve block does. This
is what the reference's StepProfile::solver times. A contact constraint-iteration is one contact
point's normal + friction update in one velocity pass: per step, substeps x 2 passes x (active
points + impact points).

Also reported: the roofline of the solve-row kernels (algorithmic bytes per launch / event-timed
launch duration, against 8 TB/s), measured on `--profile-steps` further steps after the timed ones
(the events around each launch slow a step down, so the timed steps run without them), and a
single-threaded CPU baseline on the same pile state: the reference engine built from its sources
where that build is present, else the CPU oracle.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from twofold_amd import scenes  # noqa: E402
from twofold_amd.solver import Solver, device_synchronize  # noqa: E402
from twofold_amd.world import HostWorld, Simulation  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Memory-side bytes per k_pass launch on this workload, from two rocprofv3 --pmc runs of this script
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per dispatch, the factors from a 256 MiB copy measured in the same
# runs, --copy-mib 256); counters cannot be read from inside the process, so the figure is carried here.
K_PASS_TRAFFIC_BYTES = 102.4e6
K_PASS_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, mean of the last 32 k_pass dispatches, scaled by a "
                         "256 MiB copy (profiles/k_pass_pile32k_traffic.md)")
# The same for the batched worlds' passes (mean of the 16 batched k_pass dispatches of 2 timed steps) and
# for the adjoint's reverse passes (mean of the last 32 k_pass_adj dispatches), measured the same way.
BATCHED_TRAFFIC_BYTES = 2.815e9
BATCHED_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, mean of the batched k_pass dispatches of 2 timed "
                          "steps, scaled by a 256 MiB copy (profiles/k_pass_batched4096_kernel_stats.md)")
ADJOINT_TRAFFIC_BYTES = 868.6e6
ADJOINT_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, mean of the last 32 k_pass_adj dispatches, scaled by "
                          "a 256 MiB copy (profiles/k_pass_adj_adjoint1024_kernel_stats.md)")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def iterations(counts, substeps):
    return substeps * 2 * (counts["points"] + counts["impact_points"])


def _one_core():
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)
    except OSError:
        pass
    try:
        os.sched_setaffinity(0, {min(os.sched_getaffinity(0))})
    except (AttributeError, OSError):
        pass


def cpu_reference(scene, bodies, substeps, budget_s):
    """The reference engine itself (oracle/_ref/libtwofold_ref.so, built from its sources), one thread:
    the same pile put in the same body state, stepped with its own narrowphase, its solve block timed by
    its own StepProfile::solver (src/world.cpp:1644-1692)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refworld

    if not refworld.ref_available():
        return None
    _one_core()
    w = refworld.RefWorld(scene)
    w.set_bodies(bodies)
    for _ in range(2):  # contacts and their warm-start impulses build up
        w.step(record=False)
    steps, its, spent = 0, 0, 0.0
    while steps < 2 or spent < budget_s:
        w.step(record=False)
        spent += w.profile()["solver"] * 1e-3
        c = w.counts()
        its += substeps * 2 * int(c[1] + c[3])
        steps += 1
    return dict(value=its / spent, unit="iterations/s", cores=1, kind="reference",
                sample=f"{steps} steps of the reference on the same settled pile from the same body state "
                       f"(its own narrowphase), solve block timed by StepProfile::solver: {spent:.1f} s of solve, "
                       f"one thread, CPU: {cpu_model()}", ms_per_step=1e3 * spent / steps)


def cpu_reference_worlds(worlds, sample, states, timed_steps=4):
    """The reference engine (oracle/_ref) stepping each sampled world alone on one thread, from the
    batch's state of that world (2 untimed steps, then `timed_steps` timed by its own
    StepProfile::solver): world-steps/s with every world weighted alike, as the batch steps them."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refworld

    if not refworld.ref_available():
        return None
    _one_core()
    per_world = []
    for k in sample:
        w = refworld.RefWorld(worlds[k])
        w.set_bodies(states[k])
        for _ in range(2):
            w.step(record=False)
        t = 0.0
        for _ in range(timed_steps):
            w.step(record=False)
            t += w.profile()["solver"] * 1e-3
        per_world.append(t / timed_steps)
    mean = sum(per_world) / len(per_world)
    return dict(value=1.0 / mean, unit="world-steps/s", cores=1, kind="reference",
                sample=f"{len(sample)} of the worlds, each alone from the batch's state of it, {timed_steps} steps each "
                       f"of the reference (its own narrowphase), solve block timed by StepProfile::solver: "
                       f"{1e3 * sum(per_world) * timed_steps:.0f} ms of solve, one thread, CPU: {cpu_model()}",
                ms_per_world_step=1e3 * mean)


def cpu_port(sim, bodies, contacts, budget_s):
    """The oracle (oracle/liboracle.so, a single-threaded CPU restatement of the reference solver, pinned
    bit-exact to it) timed on the same input, with the same refresh callback run on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refworld

    _one_core()
    lib = refworld._load_oracle()
    b = np.ascontiguousarray(bodies.copy())
    c = np.ascontiguousarray(contacts.copy())
    n = max(len(c), 1)
    arrays = [np.zeros(n, np.uint32) for _ in range(6)]
    colors = np.zeros(n, np.uint8)
    L = refworld._Lists()
    for name, a in zip(("active", "fast", "fast_solid", "impact", "by_color", "impact_by_color"), arrays):
        setattr(L, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    L.colors = colors.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    fn = ctypes.cast(sim.world.refresh_fn, refworld.REFRESH_FN)
    steps, its, spent = 0, 0, 0.0
    substeps = int(sim.settings["substeps"])
    while steps == 0 or spent < budget_s:
        t0 = time.perf_counter()
        rc = lib.tfo_step(b.ctypes.data, len(b), c.ctypes.data, len(c), sim.settings.ctypes.data, np.float32(sim.dt),
                          fn, ctypes.c_void_p(sim.world.h), ctypes.byref(L))
        spent += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")
        its += substeps * 2 * (L.points_active + L.points_impact)
        steps += 1
    return dict(value=its / spent, unit="iterations/s", cores=1, kind="port",
                sample=f"{steps} solve steps of the same settled pile from the same input, oracle/liboracle.so "
                       f"(-O2, one thread, refresh narrowphase on the same thread), {spent:.1f} s, CPU: {cpu_model()}",
                ms_per_step=1e3 * spent / steps)


def device_sync():
    """hipDeviceSynchronize through the library's own HIP runtime."""
    return device_synchronize


class SolveSteps:
    """Steps a Simulation the way the benchmark times it: host narrowphase and upload (untimed), the
    solve step on resident inputs (timed, bracketed by device syncs), then the download (timed apart
    with the upload, for the rate with PCIe transfers included)."""

    def __init__(self, sim):
        self.sim = sim
        self.sync = device_sync()
        self.transfer = 0.0

    def __call__(self):
        sim = self.sim
        if sim.stream and sim.world.worlds == 1:
            return self.streamed()
        sim.bodies = sim.solver.bodies()
        contacts, woke = sim.world.collide(sim.bodies, sim.dt, sim.contacts)
        t = time.perf_counter()
        if woke:
            sim.upload(sim.bodies, contacts)
        else:
            sim.solver.upload_contacts(contacts)
        self.transfer += time.perf_counter() - t
        self.sync()
        t = time.perf_counter()
        sim.solver.step_native(sim.dt, sim.world.refresh_fn, sim.world.h)
        self.sync()
        spent = time.perf_counter() - t
        t = time.perf_counter()
        sim.contacts = sim.solver.contacts()
        sim.bodies = sim.solver.bodies()
        self.transfer += time.perf_counter() - t
        sim.steps += 1
        return spent

    def streamed(self):
        """A step with the contacts resident (tfs_stream_contacts): the bodies come back (timed with the
        transfers), the host finds the candidate pairs and their manifolds (untimed, as the narrowphase),
        the pairs and manifolds go over and the device updates its contacts (timed with the transfers),
        then the solve (timed)."""
        sim = self.sim
        t = time.perf_counter()
        poses = sim.solver.poses()
        self.transfer += time.perf_counter() - t
        for f in ("p", "q", "com", "v", "w", "awake"):
            sim.bodies[f] = poses[f]
        pairs, manifolds, woke = sim.world.pairs(sim.bodies, sim.dt, sim.pinned)
        t = time.perf_counter()
        sim.solver.stream(pairs, manifolds)
        if woke:
            # the whole bodies, woken where the pairs woke them (the rest of the host copy is stale)
            woken = sim.bodies["awake"].astype(bool) & ~poses["awake"].astype(bool)
            full = sim.solver.bodies()
            full["awake"][woken] = 1
            full["sleep_time"][woken] = 0.0
            sim.solver.upload_bodies(full)
            sim.bodies = full
        self.sync()
        self.transfer += time.perf_counter() - t
        t = time.perf_counter()
        sim.solver.step_native(sim.dt, sim.world.refresh_fn, sim.world.h)
        self.sync()
        spent = time.perf_counter() - t
        sim.contacts = range(sim.solver.nc)   # only their number is read
        sim.steps += 1
        return spent


def tune(solver, args, k):
    """Starts the pass-shape autotuning at the settle step that lets it end with the settle (tuning
    candidates x rounds steps; about 36 at most)."""
    if args.tune_rounds > 0 and k == max(0, args.settle - 18 * args.tune_rounds):
        solver.autotune(args.tune_rounds)


def finish_tuning(solver, step):
    """Steps on until tuning is done (when the settle was shorter than it; at most 200 more steps),
    then the shape kept."""
    for _ in range(200):
        if solver.pass_shape()["tuning_steps_left"] == 0:
            break
        step()
    solver.autotune(0)
    x = solver.pass_shape()
    return {"blocks": x["blocks"], "tail_rows": x["tail_rows"], "made_axes": x["made_axes"], "fused_substep": x["fused"]}


def run_batched(args):
    """Batched worlds: `--worlds` independent piles of log-uniform [8, 1024] bodies (the pile recipe
    per world, its own friction and restitution), solved together in one resident set
    (tfs_set_worlds), timed per solve step as the pile is; against looping the plain path over a
    sample of the same worlds, each alone in its own solve step from the same state."""
    sizes = scenes.batch_sizes(args.worlds)
    t0 = time.perf_counter()
    worlds = [scenes.batch_world(n, 0x5EED0000 + k) for k, n in enumerate(sizes)]
    sim = Simulation(worlds, device_refresh=args.refresh == "device")
    try:
        return _run_batched(args, worlds, sim, t0)
    finally:
        sim.close()


def _run_batched(args, worlds, sim, t0):
    log(f"{len(worlds)} worlds, {len(sim.bodies)} bodies, built in {time.perf_counter() - t0:.1f} s; settling "
        f"{args.settle} steps")
    step = SolveSteps(sim)
    for k in range(args.settle):
        tune(sim.solver, args, k)
        step()
        if (k + 1) % 10 == 0:
            log(f"settle step {k + 1}: {len(sim.contacts)} contacts, {time.perf_counter() - t0:.1f} s")
    shape = finish_tuning(sim.solver, step)
    # the substep loop of each step as one HIP graph (the refreshes collide on the device)
    sim.solver.set_graph(args.graph)
    for _ in range(args.warmup):
        step()
    wall = 0.0
    contacts = 0
    for _ in range(args.steps):
        wall += step()
        contacts += len(sim.contacts)
    batched = len(worlds) * args.steps / wall
    log(f"batched: {args.steps} steps in {wall:.3f} s of solve, {batched:.4g} world-steps/s (graph {args.graph}, "
        f"shape {shape})")
    # the other launch forms, for comparison: the tuned shape with the other graph setting, and the
    # fused substep kernel in a graph
    sim.solver.set_graph(not args.graph)
    other = sum(step() for _ in range(args.compare_steps)) / max(args.compare_steps, 1)
    tuned = sim.solver.pass_shape()
    sim.solver.set_pass_shape(min(tuned["blocks"], 128), 64, tuned["made_axes"], True)
    sim.solver.set_graph(True)
    fused_graph = sum(step() for _ in range(args.compare_steps)) / max(args.compare_steps, 1)
    sim.solver.set_pass_shape(tuned["blocks"], tuned["tail_rows"], tuned["made_axes"], tuned["fused"])
    sim.solver.set_graph(args.graph)
    log(f"batched with graph {not args.graph}: {1e3 * other:.3f} ms per step; fused substeps in a graph: "
        f"{1e3 * fused_graph:.3f} ms per step")
    sim.solver.set_profiling(True)
    for _ in range(args.batch_profile_steps):
        step()
    sim.solver.set_profiling(False)
    launches, ms, algo = sim.solver.profile(1)
    _, wait_ms, _ = sim.solver.profile(3)
    refreshes, narrow_ms, pairs = sim.solver.profile(4)
    family_ms = [sim.solver.profile(f)[1] / max(args.batch_profile_steps, 1) for f in range(6)]
    achieved = algo / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    log(f"solve passes: {launches} launches, {ms:.2f} ms, {algo / max(launches, 1) / 1e6:.2f} MB per launch; "
        f"ms per step by profile family: {', '.join(f'{x:.2f}' for x in family_ms)}")

    # looping: the plain path on each sampled world alone, from the batch's current state
    rng = scenes.Rng(0x1005)
    sample = sorted({int(rng.next() * len(worlds)) % len(worlds) for _ in range(args.loop_sample)})
    sync = device_sync()
    solver = Solver(64, 256, settings=sim.settings)
    per_world = []
    try:
        for k in sample:
            lo, hi = sim.world.body_start[k], sim.world.body_start[k + 1]
            host = HostWorld(worlds[k], sim.settings)
            bodies = np.ascontiguousarray(sim.bodies[lo:hi])
            contacts_k, _ = host.collide(bodies.copy(), sim.dt)
            solver.set_body_shapes(None)
            for timed in (False, True):
                solver.upload(bodies, contacts_k)
                if sim.device_refresh is not None:
                    solver.set_body_shapes(host.body_shapes())
                sync()
                t = time.perf_counter()
                solver.step_native(sim.dt, host.refresh_fn, host.h)
                sync()
                if timed:
                    per_world.append(time.perf_counter() - t)
            host.close()
    finally:
        solver.close()
    looping = 1.0 / (sum(per_world) / len(per_world))
    log(f"looping the plain path over {len(sample)} worlds: {1e3 * sum(per_world) / len(per_world):.3f} ms per world-step")
    cpu = None
    if not args.no_cpu:
        states = {k: np.ascontiguousarray(sim.bodies[sim.world.body_start[k]:sim.world.body_start[k + 1]]) for k in sample}
        cpu = cpu_reference_worlds(worlds, sample, states)
        if cpu is not None:
            log(f"cpu baseline (reference): {cpu['value']:.4g} world-steps/s")
    return {
        "metric": "batched world-steps/s",
        "value": batched,
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded per-world pile recipes; host narrowphase between steps, mid-step refreshes "
                f"collided on the {'device' if sim.device_refresh is not None else 'host'}",
        "config": {
            "workload": f"{len(worlds)} independent worlds of log-uniform [8, 1024] bodies ({len(sim.bodies)} "
                        f"bodies), after {args.settle} drop steps, dt 1/60, 4 substeps; per step: host narrowphase "
                        f"+ upload (untimed), then one solve block + sleep pass over all worlds (timed; its mid-step "
                        f"refreshes collide on the {'device' if sim.device_refresh is not None else 'host'})",
            "worlds": len(worlds), "bodies": len(sim.bodies), "contacts_per_step": contacts / args.steps,
            "looping_world_steps_per_s": looping, "looping_sample_worlds": len(sample),
            "batched_over_looping": batched / looping, "pass_shape": shape,
            "graph": bool(args.graph),
            ("ms_per_step_without_graph" if args.graph else "ms_per_step_with_graph"): 1e3 * other,
            "ms_per_step_fused_substeps_in_graph": 1e3 * fused_graph,
            "launch_forms_note": "the tuned shape (the autotuner's candidates include the fused substep kernel) "
                                 "without and with the per-step HIP graph of the substep loop, and the fused "
                                 "substep kernel in a graph; each timed on its own steps of the same run",
            "refreshes_per_step": refreshes / max(args.batch_profile_steps, 1),
            "refresh_pairs_per_step": pairs / max(args.batch_profile_steps, 1),
            "refresh_wait_ms_per_step": wait_ms / max(args.batch_profile_steps, 1),
            "refresh_narrowphase_ms_per_step": narrow_ms / max(args.batch_profile_steps, 1),
            "refresh": "device" if sim.device_refresh is not None else "host callback",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": BATCHED_TRAFFIC_BYTES if args.worlds == 4096 else None,
            "traffic_source": BATCHED_TRAFFIC_SOURCE if args.worlds == 4096 else None,
            "kernel": "k_pass<bias|relax> (one grid-barrier launch per solve pass over all worlds' rows)",
            "launches": launches, "profiled_steps": args.batch_profile_steps,
            "bytes_per_launch": algo / max(launches, 1), "ms_per_launch": ms / max(launches, 1),
        },
        "cpu_baseline": cpu,
    }


def run_adjoint(args):
    """The fixed-contact model and its adjoint on `--worlds` throw worlds (64 boxes thrown at a 10-box
    wall each; twofold_amd.scenes.throw_world): `--steps` forward steps (each checkpointed), then the
    reverse sweep of every step from its checkpoint (tfs_adjoint_step: the step again with its tape
    recorded, then backwards through it), seeded with d loss / d com of the squared distance of the
    final centres of mass to seeded targets. Both legs are timed per step, device-synchronised; the
    headline is adjoint time / forward time."""
    t0 = time.perf_counter()
    worlds = [scenes.throw_world(0xAD000000 + k, steps=args.steps) for k in range(args.worlds)]
    host = HostWorld(worlds)
    bodies = host.bodies()
    contacts, _ = host.collide(bodies, host.dt)
    body_start = np.asarray(host.body_start, np.uint32)
    moving = bodies["type"] == 2
    target = bodies["com"].astype(np.float64) + np.random.default_rng(0xAD).uniform(-0.5, 0.5, (len(bodies), 3))
    log(f"{len(worlds)} worlds, {len(bodies)} bodies, {len(contacts)} contacts, built in {time.perf_counter() - t0:.1f} s")
    sync = device_sync()
    solver = Solver(len(bodies), len(contacts), settings=host.settings)
    try:
        for rep in range(args.warmup + 1):
            solver.upload_worlds(bodies, body_start, contacts)
            solver.fixed_begin(host.dt, args.steps)
            sync()
            t = time.perf_counter()
            for _ in range(args.steps):
                solver.fixed_step()
            sync()
            forward = time.perf_counter() - t
            com = solver.bodies()["com"].astype(np.float64)
            dl = np.zeros_like(com)
            dl[moving] = 2.0 * (com[moving] - target[moving])
            loss = float(np.sum((com[moving] - target[moving]) ** 2))
            solver.adjoint_seed(dl)
            sync()
            t = time.perf_counter()
            for _ in range(args.steps):
                solver.adjoint_step()
            sync()
            backward = time.perf_counter() - t
            gv, gw, gmu = solver.adjoint_read()
            log(f"round {rep}: forward {1e3 * forward / args.steps:.3f} ms/step, adjoint {1e3 * backward / args.steps:.3f} "
                f"ms/step, loss {loss:.6g}, |dL/dv0| {np.linalg.norm(gv):.6g}")
        counts = solver.counts()
        # one more round event-timed: the reverse sweep's passes (profile family 6) for the roofline
        solver.upload_worlds(bodies, body_start, contacts)
        solver.fixed_begin(host.dt, args.steps)
        for _ in range(args.steps):
            solver.fixed_step()
        solver.adjoint_seed(dl)
        solver.set_profiling(True)
        for _ in range(args.steps):
            solver.adjoint_step()
        solver.set_profiling(False)
        launches, ms, algo = solver.profile(6)
        achieved = algo / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        log(f"reverse passes: {launches} launches, {ms:.2f} ms, {algo / max(launches, 1) / 1e6:.2f} MB per launch")
    finally:
        solver.close()
        host.close()
    ratio = backward / forward
    cpu = None
    if not args.no_cpu:
        cpu = cpu_adjoint(worlds[:args.cpu_worlds], args)
    return {
        "metric": "adjoint/forward time",
        "value": ratio,
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * backward / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded throw worlds (64 boxes with N(0, 2 m/s), N(0, 1 rad/s) at a 10-box wall)",
        "config": {
            "workload": f"{len(worlds)} worlds x 74 boxes ({len(bodies)} bodies, {len(contacts)} contacts fixed at the "
                        f"first step), {args.steps} steps of dt 1/60, 4 substeps; forward = checkpoint + step, adjoint = "
                        f"step again recording + reverse sweep, per step",
            "worlds": len(worlds), "bodies": len(bodies), "contacts": len(contacts),
            "forward_ms_per_step": 1e3 * forward / args.steps, "adjoint_ms_per_step": 1e3 * backward / args.steps,
            "forward_world_steps_per_s": len(worlds) * args.steps / forward,
            "adjoint_world_steps_per_s": len(worlds) * args.steps / backward,
            "colors": counts["colors"], "points": counts["points"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": ADJOINT_TRAFFIC_BYTES if args.worlds == 1024 else None,
            "traffic_source": ADJOINT_TRAFFIC_SOURCE if args.worlds == 1024 else None,
            "kernel": "k_pass_adj<bias|relax> (one grid-barrier launch per reverse pass over all worlds' rows; bytes "
                      "per row 440 + 348n)",
            "launches": launches, "profiled_steps": args.steps,
            "bytes_per_launch": algo / max(launches, 1), "ms_per_launch": ms / max(launches, 1),
        },
        "cpu_baseline": cpu,
    }


def cpu_adjoint(worlds, args):
    """The float64 model of oracle/fixed_contacts.py (numpy, one thread) on the first worlds of the
    workload: forward and reverse over the same steps, world-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import refworld
        from fixed_contacts import FixedContacts
    except ImportError as e:
        log(f"cpu baseline skipped: {e}")
        return None
    host = HostWorld(worlds)
    bodies = host.bodies()
    contacts, _ = host.collide(bodies, host.dt)
    lib = refworld._load_oracle()
    b = bodies.copy()
    for i in range(len(b)):
        lib.tfo_update_inverse_mass(b[i:i + 1].ctypes.data)
    colors = np.full(len(contacts), 255, np.uint8)
    lst = np.arange(len(contacts), dtype=np.uint32)
    lib.tfo_color(b.ctypes.data, len(b), contacts.ctypes.data, lst.ctypes.data, len(lst), colors.ctypes.data)
    model = FixedContacts(bodies, contacts, colors, host.settings, float(host.dt))
    steps = args.cpu_steps
    t = time.perf_counter()
    model.adjoint(model.com0, steps=steps)
    spent = time.perf_counter() - t
    return {"value": len(worlds) * steps / spent, "unit": "world-steps/s (forward + reverse)", "cores": 1,
            "kind": "port", "cpu": cpu_model(),
            "sample": f"{len(worlds)} worlds x {steps} steps of the float64 numpy model (oracle/fixed_contacts.py)"}


def run_pile(args, planar=False):
    """The 32,768-body pile (the headline) or, with `planar`, the planar pile: drop and settle with every
    step's solve timed (the whole-run rate), then the timed window, then (pile) a leg with the mid-step
    refreshes through the host callback, then event-timed steps for the roofline, then the CPU baseline
    on the state at the end of the settle."""
    scene = scenes.planar_pile() if planar else scenes.pile(side=args.side)
    t0 = time.perf_counter()
    # the planar pile's prisms and cylinders are collided on the host only
    sim = Simulation(scene, device_refresh=args.refresh == "device" and not planar)
    log(f"{scene.name}: {len(sim.bodies)} bodies built in {time.perf_counter() - t0:.1f} s; settling {args.settle} steps")
    substeps = int(sim.settings["substeps"])
    solve_step = SolveSteps(sim)
    run = dict(its=0, wall=0.0, steps=0)

    def whole(spent):
        run["its"] += iterations(sim.solver.counts(), substeps)
        run["wall"] += spent
        run["steps"] += 1

    try:
        for k in range(args.settle):
            tune(sim.solver, args, k)
            whole(solve_step())
            if (k + 1) % args.log_every == 0:
                c = sim.solver.counts()
                v = np.linalg.norm(sim.bodies["v"], axis=1)
                log(f"settle step {k + 1}: {len(sim.contacts)} contacts, {c['points']} points, {c['colors']} colours, "
                    f"{c['rest']} rest, {c['rest_levels']} rest levels, speed mean {v.mean():.3f} max {v.max():.2f} m/s, "
                    f"{time.perf_counter() - t0:.1f} s")
        shape = finish_tuning(sim.solver, solve_step)
        bodies0 = sim.solver.bodies()
        for _ in range(args.warmup):
            whole(solve_step())
        solve_step.transfer = 0.0
        its = 0
        solve_ms = 0.0
        wall = 0.0
        contacts = 0
        first = run["steps"] + 1
        for _ in range(args.steps):
            spent = solve_step()
            wall += spent
            counts = sim.solver.counts()
            its += iterations(counts, substeps)
            solve_ms += sim.solver.last_ms()[0]
            contacts += len(sim.contacts)
            whole(spent)
        window = f"{first}-{run['steps']}"
        transfers = solve_step.transfer
        while run["steps"] < args.whole_steps:
            whole(solve_step())
        host_leg = None
        if not planar and sim.device_refresh is not None and args.host_refresh_steps > 0:
            # the drop-in configuration: the same steps with the refreshes through the host callback
            sim.solver.set_body_shapes(None)
            h_its, h_wall = 0, 0.0
            for _ in range(args.host_refresh_steps):
                h_wall += solve_step()
                h_its += iterations(sim.solver.counts(), substeps)
            sim.solver.set_body_shapes(sim.device_refresh)
            host_leg = {"value": h_its / h_wall, "unit": "iterations/s", "steps": args.host_refresh_steps,
                        "ms_per_step": 1e3 * h_wall / args.host_refresh_steps,
                        "window": f"{run['steps'] + 1}-{run['steps'] + args.host_refresh_steps}",
                        "refresh": "host callback (twofold_amd/csrc/scene.cpp tfw_refresh)"}
            log(f"host-refresh leg: {host_leg['value']:.4g} iterations/s")
        streamed = None
        if args.stream_steps > 0:
            # the contacts resident on the device (tfs_stream_contacts): only the candidate pairs and their
            # manifolds go over each step, and the bodies come back; the first streamed step starts from
            # contacts that are not reusable (they were uploaded), so it and the warm-up are not timed
            sim.stream = True
            for _ in range(args.warmup):
                solve_step()
            solve_step.transfer = 0.0
            s_its, s_wall, first = 0, 0.0, run["steps"] + args.warmup + 1
            for _ in range(args.stream_steps):
                s_wall += solve_step()
                s_its += iterations(sim.solver.counts(), substeps)
            s_transfer = solve_step.transfer
            c = sim.solver.counts()
            streamed = {"value": s_its / s_wall, "unit": "iterations/s", "steps": args.stream_steps,
                        "window": f"{first}-{first + args.stream_steps - 1}",
                        "ms_per_step": 1e3 * s_wall / args.stream_steps,
                        "transfer_ms_per_step": 1e3 * s_transfer / args.stream_steps,
                        "value_with_transfers": s_its / (s_wall + s_transfer), "colors": c["colors"],
                        "contacts": sim.solver.nc,
                        "note": "contacts resident on the device, updated each step from the candidate pairs and "
                                "their manifolds (tfs_stream_contacts: reuse of contacts that still fit, impulse "
                                "carry); transfers = the bodies' poses back (tfs_download_poses) + the pairs and "
                                "manifolds over + the stream's device work"}
            log(f"streamed: {streamed['value']:.4g} iterations/s, {streamed['value_with_transfers']:.4g} with transfers "
                f"({streamed['transfer_ms_per_step']:.2f} ms of transfers per step)")
            sim.stream = False
            sim.contacts = sim.solver.contacts()
        sim.solver.set_profiling(True)
        for _ in range(args.profile_steps):
            solve_step()
        sim.solver.set_profiling(False)
        # the dominant kernel: the coloured passes on the pile, the rest passes (planar contacts) on the planar pile
        launches, ms, algo = sim.solver.profile(5 if planar else 1)
        _, wait_ms, _ = sim.solver.profile(3)
        refreshes, narrow_ms, pairs = sim.solver.profile(4)
        achieved = algo / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        log(f"timed {args.steps} solve steps: {wall:.3f} s; {'rest' if planar else 'solve'} passes: {launches} launches, "
            f"{ms:.2f} ms, {algo / max(launches, 1) / 1e6:.2f} MB per launch")
        cpu = None
        if not args.no_cpu:
            cpu = cpu_reference(scene, bodies0, substeps, args.cpu_seconds)
            if cpu is None:
                cpu = cpu_port(sim, bodies0, sim.world.collide(bodies0.copy(), sim.dt, None)[0], args.cpu_seconds)
            log(f"cpu baseline ({cpu['kind']}): {cpu['value']:.3e} iterations/s ({cpu['ms_per_step']:.0f} ms/step)")
        refresh_where = "device" if sim.device_refresh is not None else "host"
    finally:
        sim.close()
    traffic = pile_traffic(args) if not planar else None
    return {
        "metric": "contact constraint-iterations/s",
        "value": its / wall,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded {scene.name} recipe; host narrowphase between steps, mid-step refreshes "
                f"collided on the {'device (tfs_set_body_shapes)' if refresh_where == 'device' else 'host (refresh callback)'}",
        "config": {
            "workload": (f"{len(bodies0)} bodies: 65,536 bodies locked to the xy plane (40% boxes, 40% prisms of 3-8 "
                         f"sides, 20% cylinders) and 4,096 free spheres in a U container" if planar else
                         f"{len(bodies0)}-body box/sphere pile in walls")
                        + f", timed over steps {window} after {args.settle} drop-and-settle steps, dt 1/60, {substeps} "
                        f"substeps; per step: host narrowphase + upload (untimed), then one solve block + sleep pass "
                        f"on the GPU (timed; its mid-step refreshes collide on the {refresh_where})",
            "timed_steps": window,
            "bodies": len(bodies0), "contacts_per_step": contacts / args.steps, "active_points": counts["points"],
            "impact_points": counts["impact_points"], "colors": counts["colors"], "rest": counts["rest"],
            "rest_levels": counts["rest_levels"], "impact_rest_levels": counts["impact_rest_levels"],
            "world_steps_per_s": args.steps / wall, "gpu_solve_ms_per_step": solve_ms / args.steps,
            "transfer_ms_per_step": 1e3 * transfers / args.steps,
            "value_with_transfers": its / (wall + transfers), "pass_shape": shape,
            "contacts": "uploaded from the host each step (tfs_upload_contacts); streamed form below",
            "streamed": streamed,
            "whole_run": {"steps": f"1-{run['steps']}", "value": run["its"] / run["wall"], "unit": "iterations/s",
                          "ms_per_step": 1e3 * run["wall"] / run["steps"],
                          "note": "every step's solve timed the same way from the first drop step on, the pass-shape "
                                  "tuning steps included"},
            "value_host_refresh": host_leg,
            "refreshes_per_step": refreshes / max(args.profile_steps, 1),
            "refresh_pairs_per_step": pairs / max(args.profile_steps, 1),
            "refresh_wait_ms_per_step": wait_ms / max(args.profile_steps, 1),
            "refresh_narrowphase_ms_per_step": narrow_ms / max(args.profile_steps, 1),
            "refresh": "device" if refresh_where == "device" else "host callback",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic[0] if traffic else None, "traffic_source": traffic[1] if traffic else None,
            "kernel": ("k_rest_walk<bias|relax> (one solve pass over the active rest list: k_rest_pre, the leading "
                       "wide levels in a grid-barrier k_rest_levels, the narrow rest walked by k_rest_walk in one "
                       "work-group; bytes per planar contact 248 + 68n)" if planar
                       else "k_pass<bias|relax> (one grid-barrier launch per solve pass)"),
            "launches": launches, "profiled_steps": args.profile_steps,
            "bytes_per_launch": algo / max(launches, 1), "ms_per_launch": ms / max(launches, 1),
        },
        "cpu_baseline": cpu,
    }


def pile_traffic(args):
    """The k_pass traffic figure of the 32k pile (carried from a rocprofv3 run), or None for other sizes."""
    if args.side != 32:
        return None
    return K_PASS_TRAFFIC_BYTES, K_PASS_TRAFFIC_SOURCE


DEFAULTS = {  # steps, warmup, settle per workload when it runs alone or as a part of the full run
    "pile": (60, 5, 240),
    "planar": (20, 2, 720),
    "batched": (10, 2, 60),
    "adjoint": (32, 2, 0),
}


def sub_args(args, workload):
    """The arguments of a workload run as a part of the full run: its own defaults."""
    a = argparse.Namespace(**vars(args))
    a.workload = workload
    a.steps, a.warmup, a.settle = DEFAULTS[workload]
    a.worlds = 1024 if workload == "adjoint" else 4096
    a.cpu_seconds = args.sub_cpu_seconds
    return a


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", choices=("all", "pile", "batched", "planar", "adjoint"), default="all",
                    help="all (default): the 32k-body pile (the headline) with the planar, batched and adjoint "
                         "workloads in sub-objects of the same line; or one of them alone")
    ap.add_argument("--worlds", type=int, default=None, help="batched: number of worlds (4096); adjoint (1024)")
    ap.add_argument("--cpu-worlds", type=int, default=4, help="adjoint: worlds of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=4, help="adjoint: steps of the CPU baseline sample")
    ap.add_argument("--loop-sample", type=int, default=64, help="batched: worlds timed alone for the looping rate")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (pile 60, planar 20, batched 10, adjoint 32)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before them (pile 5, others 2)")
    ap.add_argument("--settle", type=int, default=None,
                    help="steps of drop-and-settle before the warm-up (pile 240, planar 720, batched 60)")
    ap.add_argument("--whole-steps", type=int, default=300, help="pile: steps of the whole-run rate (1 to this)")
    ap.add_argument("--host-refresh-steps", type=int, default=10,
                    help="pile: steps timed with the refreshes through the host callback (the drop-in form)")
    ap.add_argument("--side", type=int, default=32, help="bodies per side of the grid (32 -> 32,768)")
    ap.add_argument("--profile-steps", type=int, default=10, help="event-timed steps for the roofline")
    ap.add_argument("--batch-profile-steps", type=int, default=3, help="batched: event-timed steps for the roofline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the headline's CPU baseline leg")
    ap.add_argument("--sub-cpu-seconds", type=float, default=8.0, help="budget of the other workloads' CPU legs")
    ap.add_argument("--time-budget", type=float, default=420.0,
                    help="all: seconds after which the workloads not yet run are skipped (noted in the line)")
    ap.add_argument("--log-every", type=int, default=20, help="settle steps between progress lines")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stream-steps", type=int, default=20,
                    help="pile and planar: steps timed after the window with the contacts resident on the device, the "
                         "candidate pairs and their manifolds sent each step (tfs_stream_contacts)")
    ap.add_argument("--graph", type=int, default=0,
                    help="batched: launch each step's substep loop as one HIP graph (1/0; the other form is timed too)")
    ap.add_argument("--compare-steps", type=int, default=4, help="batched: steps timed with the other launch form")
    ap.add_argument("--refresh", choices=("device", "host"), default="device",
                    help="where mid-step refreshes collide their pairs: on the device (tfs_set_body_shapes; "
                         "spheres, boxes and half-spaces) or through the host world's callback")
    ap.add_argument("--tune-rounds", type=int, default=2,
                    help="rounds of pass-shape autotuning over the last settle steps (0: the default shape)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB: the reference reading for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run")
    args = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py runs on one GPU only: this build measures the solve path on a single MI355X "
              "(--gpus 1); multi-GPU runs are not part of it.", file=sys.stderr)
        sys.exit(2)
    t_start = time.perf_counter()
    one = args.workload if args.workload != "all" else "pile"
    d_steps, d_warmup, d_settle = DEFAULTS[one]
    args.worlds = args.worlds if args.worlds is not None else (1024 if one == "adjoint" else 4096)
    args.steps = args.steps if args.steps is not None else d_steps
    args.warmup = args.warmup if args.warmup is not None else d_warmup
    args.settle = args.settle if args.settle is not None else d_settle

    if args.copy_mib > 0:
        copy_reference(args.copy_mib)

    if one == "batched":
        out = run_batched(args)
    elif one == "adjoint":
        out = run_adjoint(args)
    else:
        out = run_pile(args, planar=one == "planar")
    if args.workload == "all":
        for name, fn in (("adjoint", run_adjoint), ("planar", lambda a: run_pile(a, planar=True)),
                         ("batched", run_batched)):
            spent = time.perf_counter() - t_start
            if spent > args.time_budget:
                out[name] = {"skipped": f"the run had spent {spent:.0f} s of its {args.time_budget:.0f} s budget"}
                log(f"{name}: skipped ({spent:.0f} s spent)")
                continue
            log(f"---- {name} ({spent:.0f} s into the run)")
            sub = fn(sub_args(args, name))
            for k in ("n_gpus", "scaling", "vs_baseline", "dtype"):
                sub.pop(k, None)
            out[name] = sub
        out["config"]["bench_seconds"] = time.perf_counter() - t_start
    print(json.dumps(out), flush=True)


def copy_reference(mib):
    """One device-to-device copy of `mib` MiB (hipMemcpy through the library's HIP runtime), the known
    size that the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run are scaled by."""
    from twofold_amd.solver import DeviceArray

    a = DeviceArray(np.ones(mib << 18, np.float32))
    b = DeviceArray(np.zeros(mib << 18, np.float32))
    rc = a.hip.hipMemcpy(b.ptr, a.ptr, mib << 20, 3)
    device_synchronize()
    a.close()
    b.close()
    log(f"copied {mib} MiB device to device ({rc})")


if __name__ == "__main__":
    main()
