"""FP64 DEM step throughput on one MI355X (particle-steps/s).

Workload: pure DEM, 1,000,000 spheres, radius lognormal (median 1 mm,
sigma 0.2, clipped to [0.5, 2] mm), density 2500 kg/m^3, kN = E * median
radius with E = 1e7 Pa, mu 0.5, restitution 0.5, in a 200 x 200 mm hopper
(floor + 4 walls inclined 10 degrees), gravity -9.81 z, dt = 5e-7 s, grid
cell = 2 x max radius.  The headline value is the settled state (a bed
built by sequential deposition under gravity, each sphere rolled into a
hollow of the bed, resting contacts preloaded by 0.2 um; about 5.6 stored
contacts per sphere at the start, 4.5-5 while the bed sags under its weight
over the first thousand steps; --packing picks it), averaged over 1,000
steps; the loose seeded-insertion state (no contacts yet) is measured beside
it.  A step is the reference's pure-DEM
loop body: advanceParticles, computeForces, finishParticleStep, all state
resident in HBM, run through dem_run_steps in the mode its autotuner picks
(candidate lists with eager launches, or the full scan replayed from
captured HIP graphs; the JSON line reports which ran and how many list
rebuilds the timed window held).

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (the per-contact law k_pairs, timed with HIP events on the solver's
stream), the per-step algorithmic-byte roofline, and a CPU baseline: the
reference's own CPU binary (built from its sources) on a bounded sample of
the same workload, with the oracle on 1 and on all host threads beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pairs_bytes(contacts: float, n: int) -> float:
    """Algorithmic HBM bytes of one k_pairs launch: per stored contact the
    list entry (8 B) and the history slot's 9 f64 read and written (144 B);
    per sphere its 144-byte contact record (a 128-byte kinematic line and 16
    bytes of constants, read once, as owner and as partner), its sorted-order
    and history-row indices (8 B), its list offset and count (8 B) and its
    6 f64 contact sum (48 B).  (Contacts of a row whose list straddles a
    64-entry chunk write their own 6 f64 instead; the model leaves that out.)"""
    return 152.0 * contacts + 208.0 * n


def step_bytes_per_particle(cbar: float, passes: int = 3) -> float:
    """Algorithmic bytes of a whole step per particle: 29 f64 of state read
    and written (464 B), mass/radius/inertia/id/status (48 B), cell key and
    index plus `passes` radix passes (12 + 24 P B), and per stored contact
    its history key and 9 f64 read and written (160 B)."""
    return 232.0 + 232.0 + 48.0 + 12.0 + 24.0 * passes + 160.0 * cbar


def _counter_delta(before, after):
    if before is None or after is None:
        return None
    return {k: after[k] - before[k] for k in ("evaluations", "sorts", "rebuilds", "graph_steps", "eager_steps")}


def measure_sweep(dm, scene, steps, warmup, seq_steps, count, seed, agitate=False):
    """Parameter sweep of `count` independent box scenes: one batched solver
    against the same scenes run one after another in single-scene solvers
    on the same GPU.  Every solver runs `warmup` untimed steps through
    dem_run_steps with the autotuner (candidate lists or the full scan with
    captured HIP graphs, reorder interval, pair grid), then the timed steps.
    agitate: seeded velocities U[-3, 3] m/s and spins U[-100, 100] rad/s, so
    that contacts form during the run (the scenes are loose insertions).
    Returns ms per step of each, the ratio and how the timed steps ran."""
    import time as _t

    import numpy as np

    scenes = scene.sweep_scenes(count=count, seed=seed)
    if agitate:
        rng = np.random.RandomState(seed + 100)
        for sc in scenes:
            sc.vel = rng.uniform(-3.0, 3.0, (sc.n, 3))
            sc.angvel = rng.uniform(-100.0, 100.0, (sc.n, 3))
    batch = scene.SceneBatch(scenes)
    dt = batch.combined.dt
    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    s.run_steps(warmup, dt, autotune=True)
    s.synchronize()
    c0 = s.run_counters()
    t0 = _t.perf_counter()
    s.run_steps(steps, dt)
    s.synchronize()
    batched = (_t.perf_counter() - t0) / steps
    counters = _counter_delta(c0, s.run_counters())
    s.advance(dt)
    s.compute_forces(collect=True)
    contacts = s.stats()["particle_contacts"]
    s.close()
    seq = 0.0
    seq_graph_steps = 0
    for sc in scenes:
        q = dm.CudaDemSolver()
        q.initialize(sc)
        q.compute_forces(collect=True)
        q.init_half_step(0.5 * dt)
        q.run_steps(warmup, dt, autotune=True)
        q.synchronize()
        c0 = q.run_counters()
        t0 = _t.perf_counter()
        q.run_steps(seq_steps, dt)
        q.synchronize()
        seq += (_t.perf_counter() - t0) / seq_steps
        seq_graph_steps += _counter_delta(c0, q.run_counters())["graph_steps"]
        q.close()
    return {"scenes": count, "particles": batch.n, "batched_ms_per_step": 1e3 * batched,
            "sequential_ms_per_step": 1e3 * seq, "speedup": seq / batched,
            "batched_particle_steps_per_s": batch.n / batched, "contacts_after": contacts,
            "steps": steps, "warmup_autotuned": warmup, "sequential_steps_per_scene": seq_steps,
            "batched_timed_steps": counters,
            "sequential_graph_step_share": seq_graph_steps / (count * seq_steps),
            "workload": f"{count} boxes, N ~ U[500, 20000] spheres r ~ U[1, 1.5] mm, seeded RSA, "
                        "per-scene mu/restitution/E, dt 1e-6, gravity"
                        + (", agitated (v ~ U[-3, 3] m/s, w ~ U[-100, 100] rad/s)" if agitate else ", at rest")}


def measure_adjoint(dm, scene, count, n, steps, K, seed):
    """Differentiable batch: `count` settling beds of n spheres; the
    gradient of bed height + plate load w.r.t. every scene's (mu,
    restitution, E) through `steps` steps with checkpoints every K steps.
    adjoint_over_forward = (reverse sweep incl. recomputation) / (plain
    forward run of the same steps through dem_run_steps);
    gradient_over_forward adds the checkpointed forward run to the
    numerator (the whole cost of one gradient against one plain forward
    run).  forward_python_loop_s: the plain run stepped from Python."""
    import time as _t

    import numpy as np
    from dem_mbd_cuda_amd import diff, host

    scenes = scene.settle_scenes(count=count, n=n, seed=seed)
    batch = scene.SceneBatch(scenes)
    dt = batch.combined.dt
    # the plain forward run: the multi-step path (one call, no host work
    # between steps), in the differentiable run's mode (candidate lists)
    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    s.synchronize()
    t0 = _t.perf_counter()
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    s.run_steps(steps, dt)
    s.synchronize()
    t_fwd = _t.perf_counter() - t0
    s.close()
    # the same run driven step by step from Python (the reference's loop)
    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    s.synchronize()
    t0 = _t.perf_counter()
    host.run_pure(s, batch.combined, steps=steps)
    s.synchronize()
    t_loop = _t.perf_counter() - t0
    s.close()
    q = dm.CudaDemSolver()
    q.initialize_batch(batch)
    q.synchronize()
    t0 = _t.perf_counter()
    h, f = q.diff_forward(steps, K, 4, scenes=count)
    t_dfwd = _t.perf_counter() - t0
    rng = np.random.RandomState(seed)
    t0 = _t.perf_counter()
    g = q.diff_backward(rng.normal(size=count), rng.normal(size=count))
    t_bwd = _t.perf_counter() - t0
    mem = q.device_memory_bytes()
    q.close()
    return {"scenes": count, "particles": batch.n, "steps": steps, "checkpoint_interval": K,
            "forward_s": t_fwd, "forward_python_loop_s": t_loop,
            "forward_with_checkpoints_s": t_dfwd, "backward_s": t_bwd,
            "adjoint_over_forward": t_bwd / t_fwd,
            "gradient_over_forward": (t_dfwd + t_bwd) / t_fwd, "device_bytes": mem,
            "finite_gradients": bool(np.all(np.isfinite(g))),
            "workload": f"{count} settling beds x {n} spheres r ~ U[1, 1.5] mm, 30 mm box, dt 1e-6, "
                        f"d(bed height + plate load)/d(mu, restitution, E) per scene"}


def measure_hopper(dm, scene, args, packing, steps, warmup, timing_steps):
    """Builds the 1M hopper in the given packing, runs `warmup` steps (the
    autotuner picks candidate lists / reorder interval / pair grid), times
    `steps` steps through dem_run_steps (replayed from captured HIP graphs
    in the full-scan modes, launched eagerly with candidate lists; the
    counters of the timed window say which, and how many list rebuilds it
    held), then `timing_steps` eager steps with HIP events around each phase
    (grid, scan, k_pairs, k_law, reductions)."""
    import numpy as np

    t0 = time.time()
    sc = scene.hopper_scene(n=args.n, seed=2, mesh_scale=args.mesh_scale, packing=packing)
    print(f"# {packing}: scene built in {time.time() - t0:.1f}s, {sc.n} spheres", file=sys.stderr, flush=True)
    dt = sc.dt
    s = dm.CudaDemSolver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    s.run_steps(warmup, dt, graph=True, autotune=True)
    s.synchronize()
    c0 = s.run_counters()
    t1 = time.perf_counter()
    s.run_steps(steps, dt, graph=True)
    s.synchronize()
    elapsed = time.perf_counter() - t1
    c1 = s.run_counters()
    window = _counter_delta(c0, c1)
    window["skin_m"] = c1.get("skin_pm", 0) * 1e-12  # candidate-list skin at the end of the window
    tms = []
    s.enable_timings(True)
    for _ in range(timing_steps):
        s.advance(dt)
        s.compute_forces(collect=False)
        s.finish(dt)
        tms.append(s.last_timings())
    tm = np.mean(np.array(tms), axis=0)
    s.advance(dt)
    s.compute_forces(collect=True)  # a diagnostics step for the contact count
    s.finish(dt)
    st = s.stats()
    s.close()
    out = {"sc": sc, "value": sc.n * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "phase_ms": {"grid": tm[0], "scan": tm[1], "pairs": tm[2], "law": tm[3]},
           "contacts": st["particle_contacts"], "cbar": st["particle_contacts"] / sc.n, "timed_steps": window}
    print(f"# {packing}: ms/step {out['ms_per_step']:.3f}; grid {tm[0]:.3f} ms, scan {tm[1]:.3f} ms, "
          f"k_pairs {tm[2]:.3f} ms, k_law {tm[3]:.3f} ms; contacts/particle {out['cbar']:.3f}",
          file=sys.stderr, flush=True)
    return out


def measure_coupled(dm, scene, n, steps, warmup):
    """DEM-MBD coupled mixer (n spheres in a settled-bed stand-in resting on
    the floor and walls of a fixed 4,096-triangle bin and on both faces of a
    2,048-triangle paddle in prescribed rotation, a dynamic 1,024-triangle
    plate on top): the reference's coupled loop (host body predictor and
    Newton-Euler solve between steps, 13 + 6 doubles per body through pinned
    memory each way, diagnostics every `warmup` steps) timed over `steps`
    steps after `warmup`."""
    import numpy as np

    from dem_mbd_cuda_amd import host

    t0 = time.time()
    sc = scene.mixer_scene(n=n, packing="dense")
    built = time.time() - t0
    s = dm.CudaDemSolver()
    s.initialize(sc)
    marks = {}

    def cb(step, t, extra):
        if step + 1 == warmup:
            s.synchronize()
            marks["t1"] = time.perf_counter()
        marks["contacts"] = (s.stats()["particle_contacts"], s.stats()["body_contacts"])

    host.run_coupled(s, sc, steps=warmup + steps, output_interval=warmup, on_frame=cb)
    s.synchronize()
    t2 = time.perf_counter()
    s.close()
    per = (t2 - marks["t1"]) / steps
    return {"particles": sc.n, "triangles": int(np.sum(sc.body_tri_count)), "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * per, "particle_steps_per_s": sc.n / per,
            "particle_contacts": int(marks["contacts"][0]), "body_contacts": int(marks["contacts"][1]),
            "scene_build_s": built,
            "workload": f"mixer: {sc.n} spheres r ~ U[1, 2] mm, relaxed lattice bed pressed onto a 220 mm bin "
                        "(4,096 triangles, fixed) and both faces of a paddle (2,048, prescribed omega_z = 6.28 "
                        "rad/s), plate (1,024, dynamic, 1 kg) resting on the bed, gravity, dt 1e-6, host body "
                        "solve every step"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure_cpu(dm, scene, n, steps, packing, mesh_scale):
    """CPU baseline on a bounded sample of the headline workload: a hopper of
    n spheres with the same radii, walls, packing and bed depth (the width
    scaled by sqrt(n / 1M)).
    * The reference's own CPU path: oracle/_ref/dem_mbd_cpu, the reference's
      host and legacy contact law built from its sources (ForceCaculate,
      CreatTree and the leapfrog of simulation.cpp IntegrateDem, single
      threaded by design), run on the sample written as a case directory
      with outputs off, once for 1 step and once for 1 + `steps`; the
      difference over `steps` is its time per step (reading the input, the
      initial grid and evaluation cancel out).
    * The oracle (the C restatement of the CUDA law) on the same sample: one
      step after its initial evaluation, on 1 thread and on all threads."""
    import shutil
    import subprocess
    import tempfile

    from dem_mbd_cuda_amd import _abi

    sc = scene.hopper_scene(n=n, width=0.2 * (n / 1e6) ** 0.5, seed=2, mesh_scale=mesh_scale, packing=packing)
    out = {"unit": "particle-steps/s", "cores": 1, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
           "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    exe = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_cpu")
    if os.path.exists(exe):
        work = tempfile.mkdtemp(prefix="dem_refcpu_")
        try:
            t = {}
            for k in (1, 1 + steps):
                d = scene.write_case(sc, os.path.join(work, f"case{k}"), k, outputs=False)
                t0 = time.perf_counter()
                r = subprocess.run([exe, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                                   timeout=600)
                t[k] = time.perf_counter() - t0
                if r.returncode != 0:
                    raise RuntimeError("reference CPU binary failed: " + r.stdout[-2000:])
        finally:
            shutil.rmtree(work, ignore_errors=True)
        per = (t[1 + steps] - t[1]) / steps
        out.update(value=n / per, kind="reference",
                   sample=f"the reference's CPU binary (its host and legacy law built from source, single-threaded "
                          f"by design) on a {n}-sphere hopper of the headline workload ({packing} bed, width "
                          f"scaled to keep the bed depth): {steps} steps timed as the difference of a "
                          f"{1 + steps}-step and a 1-step run ({t[1 + steps]:.1f} s and {t[1]:.1f} s), outputs off",
                   reference_s_per_step=per)
    lib = _abi.Library(os.path.join(ROOT, "oracle", "liboracle.so"), "ora_")
    port = {}
    threads_all = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    for threads in sorted({1, threads_all}):
        o = lib.solver()
        o.set_threads(threads)
        o.initialize(sc)
        o.compute_forces(collect=True)
        o.init_half_step(0.5 * sc.dt)
        c1 = time.perf_counter()
        o.advance(sc.dt)
        o.compute_forces(collect=False)
        o.finish(sc.dt)
        o.state()
        port[threads] = n / (time.perf_counter() - c1)
        o.close()
    out["port"] = {"kind": "port", "what": "the oracle (C restatement of the CUDA law) on the same sample, one step",
                   "threads_1": port[1], "threads_all": port[threads_all], "threads": threads_all}
    if "value" not in out:  # no reference build: the port on one thread is the baseline
        out.update(value=port[1], kind="port", sample=out["port"]["what"] + ", 1 thread")
    else:
        out["port"]["ratio_to_reference_1_thread"] = port[1] / out["value"]
    return out


def copy_calibration(gib=1):
    """A plain device copy of `gib` GiB (read + write), the reference kernel
    for scaling FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    a = torch.ones((gib << 27,), dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter on a copy of known size), when it is for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("workload") != workload:
        return None
    return t.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=120,
                    help="untimed steps; from 104 on they include the autotuner's trial spans")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=20, help="reference-binary steps timed for the CPU baseline")
    ap.add_argument("--cpu-n", type=int, default=100_000, help="spheres of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mesh-scale", type=float, default=2.0,
                    help="grid cell in max radii (the 1M configuration states 2)")
    ap.add_argument("--packing", choices=("settled", "dense", "rsa"), default="settled",
                    help="headline state: deposited bed at rest (settled), compressed relaxed lattice "
                         "(dense, not at rest) or loose seeded insertion (rsa)")
    ap.add_argument("--no-insertion", action="store_true", help="skip the loose-insertion figure")
    ap.add_argument("--timing-steps", type=int, default=20, help="eager steps timed per phase with HIP events")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--no-coupled", action="store_true", help="skip the DEM-MBD coupled mixer figure")
    ap.add_argument("--coupled-n", type=int, default=200_000)
    ap.add_argument("--coupled-steps", type=int, default=300)
    ap.add_argument("--only-coupled", action="store_true",
                    help="measure only the coupled mixer and print its own JSON line (for profiling it alone)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched parameter-sweep figure")
    ap.add_argument("--sweep-scenes", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=100)
    ap.add_argument("--sweep-seq-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-batch figure")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="measure only the differentiable batch and print its own JSON line (for profiling it alone)")
    ap.add_argument("--adjoint-scenes", type=int, default=64)
    ap.add_argument("--adjoint-n", type=int, default=5000)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-K", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the DEM step runs on a single device)", file=sys.stderr)
        sys.exit(2)

    import dem_mbd_cuda_amd as dm
    from dem_mbd_cuda_amd import scene

    if args.only_coupled:
        c = measure_coupled(dm, scene, args.coupled_n, args.coupled_steps, 50)
        print(json.dumps({"metric": "FP64 particle-steps/s, coupled DEM-MBD mixer", "value": c["particle_steps_per_s"],
                          "unit": "particle-steps/s", "n_gpus": 1, "steps": c["steps"], "warmup": c["warmup"],
                          "ms_per_step": c["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded radii and packing)",
                          "config": {k: v for k, v in c.items() if k not in ("ms_per_step", "particle_steps_per_s")}}))
        return
    if args.only_adjoint:
        a = measure_adjoint(dm, scene, args.adjoint_scenes, args.adjoint_n, args.adjoint_steps, args.adjoint_K, 5)
        print(json.dumps({"metric": "differentiable batch: whole gradient cost over one plain forward run",
                          "value": a["gradient_over_forward"], "unit": "x forward", "n_gpus": 1,
                          "steps": a["steps"], "warmup": 0, "ms_per_step": 1e3 * a["forward_s"] / a["steps"],
                          "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic (seeded radii and packing)", "config": a}))
        return
    if args.copy_calibration:
        copy_calibration()
    head = measure_hopper(dm, scene, args, args.packing, args.steps, args.warmup, args.timing_steps)
    sc = head["sc"]
    other = None
    if not args.no_insertion:
        packing = "rsa" if args.packing != "rsa" else "settled"
        o = measure_hopper(dm, scene, args, packing, min(args.steps, 500), args.warmup, 5)
        other = {"packing": packing, "value": o["value"], "ms_per_step": o["ms_per_step"],
                 "contacts_per_particle": o["cbar"], "phase_ms": o["phase_ms"], "steps": min(args.steps, 500)}
        del o

    state = {"settled": "settled bed (sequential deposition under gravity, contacts preloaded 0.2 um)",
             "dense": "compressed lattice layers partly relaxed (not at rest)",
             "rsa": "loose seeded random sequential insertion"}[args.packing]
    workload = (f"pure DEM hopper, {sc.n} spheres lognormal r (median 1 mm, sigma 0.2, clipped 0.5-2 mm), "
                f"floor + 4 walls inclined 10 deg, gravity, dt 5e-7, grid cell {args.mesh_scale} x max radius, "
                + state)
    pairs_ms = float(head["phase_ms"]["pairs"])
    bytes_pairs = pairs_bytes(head["contacts"], sc.n)
    achieved = bytes_pairs / (pairs_ms * 1e-3) / 1e9
    traffic = load_traffic("k_pairs", workload)
    step_b = step_bytes_per_particle(head["cbar"])
    step_gbs = head["value"] * step_b / 1e9

    coupled = None
    if not args.no_coupled:
        t0 = time.time()
        coupled = measure_coupled(dm, scene, args.coupled_n, args.coupled_steps, 50)
        print(f"# coupled: {coupled['particles']} spheres + {coupled['triangles']} triangles: "
              f"{coupled['ms_per_step']:.3f} ms/step, {coupled['particle_steps_per_s']:.3g} particle-steps/s, "
              f"{coupled['body_contacts']} body contacts ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    sweep = None
    if not args.no_sweep:
        t0 = time.time()
        sweep = measure_sweep(dm, scene, args.sweep_steps, 120, args.sweep_seq_steps, args.sweep_scenes, 4)
        sweep["with_contacts"] = measure_sweep(dm, scene, args.sweep_steps, 120, args.sweep_seq_steps,
                                               args.sweep_scenes, 4, agitate=True)
        for tag, sw in (("at rest", sweep), ("agitated", sweep["with_contacts"])):
            print(f"# sweep ({tag}): {sw['scenes']} scenes, {sw['particles']} spheres, {sw['contacts_after']} "
                  f"contacts: batched {sw['batched_ms_per_step']:.3f} ms/step, sequential "
                  f"{sw['sequential_ms_per_step']:.3f} ms/step, speedup {sw['speedup']:.1f}x; timed batch steps "
                  f"{sw['batched_timed_steps']}", file=sys.stderr, flush=True)
        print(f"# sweeps took {time.time() - t0:.0f}s", file=sys.stderr, flush=True)

    adjoint = None
    if not args.no_adjoint:
        t0 = time.time()
        adjoint = measure_adjoint(dm, scene, args.adjoint_scenes, args.adjoint_n, args.adjoint_steps,
                                  args.adjoint_K, 5)
        print(f"# adjoint: {adjoint['scenes']} x {args.adjoint_n} spheres, {adjoint['steps']} steps: forward "
              f"{adjoint['forward_s']:.2f}s, forward+checkpoints {adjoint['forward_with_checkpoints_s']:.2f}s, "
              f"reverse {adjoint['backward_s']:.2f}s, ratio {adjoint['adjoint_over_forward']:.2f}, whole gradient {adjoint['gradient_over_forward']:.2f}x forward "
              f"({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = measure_cpu(dm, scene, args.cpu_n, args.cpu_steps, args.packing, args.mesh_scale)
        print(f"# cpu: reference {cpu.get('value', 0):.3g} particle-steps/s ({cpu.get('kind')}), port "
              f"{cpu['port']['threads_1']:.3g} (1 thread) / {cpu['port']['threads_all']:.3g} "
              f"({cpu['port']['threads']} threads) ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    out = {
        "metric": "FP64 particle-steps/s @1M",
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal radii and packing)",
        "config": {"workload": workload, "particles": sc.n, "dt": sc.dt, "packing": args.packing,
                   "contacts_per_particle": head["cbar"]},
        "timed_window": head["timed_steps"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_pairs", "kernel_ms": pairs_ms, "bytes_per_launch": bytes_pairs,
                     "contacts": head["contacts"]},
        "step_roofline": {"bytes_per_particle_step": step_b, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS, "unit": "GB/s"},
        "phase_ms": head["phase_ms"],
        "insertion_state": other,
        "coupled": coupled,
        "cpu_baseline": cpu,
        "batched_sweep": sweep,
        "adjoint": adjoint,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
