"""FP64 DEM step throughput on one MI355X (particle-steps/s).

Workload: pure DEM, 1,000,000 spheres, radius lognormal (median 1 mm,
sigma 0.2, clipped to [0.5, 2] mm), density 2500 kg/m^3, kN = E * median
radius with E = 1e7 Pa, mu 0.5, restitution 0.5, in a 200 x 200 mm hopper
(floor + 4 walls inclined 10 degrees), gravity -9.81 z, dt = 5e-7 s, grid
cell = 2 x max radius.  The headline value is the settled state (a
near-touching bed, about 3.7 stored contacts per sphere; --packing picks
it), averaged over 1,000 steps; the loose seeded-insertion state (no
contacts yet) is measured beside it.  A step is the reference's pure-DEM
loop body: advanceParticles, computeForces, finishParticleStep, all state
resident in HBM, replayed from captured HIP graphs.

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (the per-contact law k_pairs, timed with HIP events on the solver's
stream), the per-step algorithmic-byte roofline, and a CPU baseline (the CPU
oracle, single thread, on a bounded sample of the same workload).
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


def measure_sweep(dm, scene, steps, warmup, seq_steps, count, seed):
    """Parameter sweep of `count` independent box scenes: one batched solver
    against the same scenes run one after another in single-scene solvers
    on the same GPU.  Both sides run their steps through dem_run_steps
    (captured HIP graphs); the sequential side is also timed with eager
    launches.  Returns ms per step of each and the ratios."""
    import time as _t

    scenes = scene.sweep_scenes(count=count, seed=seed)
    batch = scene.SceneBatch(scenes)
    dt = batch.combined.dt
    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    s.run_steps(warmup, dt)
    s.synchronize()
    t0 = _t.perf_counter()
    s.run_steps(steps, dt)
    s.synchronize()
    batched = (_t.perf_counter() - t0) / steps
    contacts = s.stats()["particle_contacts"]
    s.close()
    seq = {True: 0.0, False: 0.0}
    for sc in scenes:
        for graph in (True, False):
            q = dm.CudaDemSolver()
            q.initialize(sc)
            q.compute_forces(collect=True)
            q.init_half_step(0.5 * dt)
            q.run_steps(4, dt, graph=graph)
            q.synchronize()
            t0 = _t.perf_counter()
            q.run_steps(seq_steps, dt, graph=graph)
            q.synchronize()
            seq[graph] += (_t.perf_counter() - t0) / seq_steps
            q.close()
    return {"scenes": count, "particles": batch.n, "batched_ms_per_step": 1e3 * batched,
            "sequential_ms_per_step": 1e3 * seq[True], "speedup": seq[True] / batched,
            "sequential_eager_ms_per_step": 1e3 * seq[False], "speedup_vs_eager": seq[False] / batched,
            "batched_particle_steps_per_s": batch.n / batched, "contacts": contacts,
            "steps": steps, "sequential_steps_per_scene": seq_steps,
            "workload": f"{count} boxes, N ~ U[500, 20000] spheres r ~ U[1, 1.5] mm, seeded RSA, "
                        "per-scene mu/restitution/E, dt 1e-6, gravity"}


def measure_adjoint(dm, scene, count, n, steps, K, seed):
    """Differentiable batch: `count` settling beds of n spheres; the
    gradient of bed height + plate load w.r.t. every scene's (mu,
    restitution, E) through `steps` steps with checkpoints every K steps.
    Cost ratio = (reverse sweep incl. recomputation) / (plain forward run of
    the same steps)."""
    import time as _t

    import numpy as np
    from dem_mbd_cuda_amd import diff, host

    scenes = scene.settle_scenes(count=count, n=n, seed=seed)
    batch = scene.SceneBatch(scenes)
    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    s.synchronize()
    t0 = _t.perf_counter()
    host.run_pure(s, batch.combined, steps=steps)
    s.synchronize()
    t_fwd = _t.perf_counter() - t0
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
            "forward_s": t_fwd, "forward_with_checkpoints_s": t_dfwd, "backward_s": t_bwd,
            "adjoint_over_forward": t_bwd / t_fwd, "device_bytes": mem,
            "finite_gradients": bool(np.all(np.isfinite(g))),
            "workload": f"{count} settling beds x {n} spheres r ~ U[1, 1.5] mm, 30 mm box, dt 1e-6, "
                        f"d(bed height + plate load)/d(mu, restitution, E) per scene"}


def measure_hopper(dm, scene, args, packing, steps, warmup, timing_steps):
    """Builds the 1M hopper in the given packing, runs `warmup` steps (the
    autotuner picks candidate lists / reorder interval / pair grid), times
    `steps` graph-replayed steps, then `timing_steps` eager steps with HIP
    events around each phase (grid, scan, k_pairs, k_law, reductions)."""
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
    t1 = time.perf_counter()
    s.run_steps(steps, dt, graph=True)
    s.synchronize()
    elapsed = time.perf_counter() - t1
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
           "contacts": st["particle_contacts"], "cbar": st["particle_contacts"] / sc.n}
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
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mesh-scale", type=float, default=2.0,
                    help="grid cell in max radii (the 1M configuration states 2)")
    ap.add_argument("--packing", choices=("dense", "rsa"), default="dense",
                    help="headline state: settled near-touching bed (dense) or loose seeded insertion (rsa)")
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
    if args.copy_calibration:
        copy_calibration()
    head = measure_hopper(dm, scene, args, args.packing, args.steps, args.warmup, args.timing_steps)
    sc = head["sc"]
    other = None
    if not args.no_insertion:
        packing = "rsa" if args.packing == "dense" else "dense"
        o = measure_hopper(dm, scene, args, packing, min(args.steps, 500), args.warmup, 5)
        other = {"packing": packing, "value": o["value"], "ms_per_step": o["ms_per_step"],
                 "contacts_per_particle": o["cbar"], "phase_ms": o["phase_ms"], "steps": min(args.steps, 500)}
        del o

    state = ("settled near-touching bed (lattice layers relaxed to contact)" if args.packing == "dense"
             else "loose seeded random sequential insertion")
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
        sweep = measure_sweep(dm, scene, args.sweep_steps, 10, args.sweep_seq_steps, args.sweep_scenes, 4)
        print(f"# sweep: {sweep['scenes']} scenes, {sweep['particles']} spheres: batched "
              f"{sweep['batched_ms_per_step']:.3f} ms/step, sequential {sweep['sequential_ms_per_step']:.3f} "
              f"ms/step, speedup {sweep['speedup']:.1f}x ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    adjoint = None
    if not args.no_adjoint:
        t0 = time.time()
        adjoint = measure_adjoint(dm, scene, args.adjoint_scenes, args.adjoint_n, args.adjoint_steps,
                                  args.adjoint_K, 5)
        print(f"# adjoint: {adjoint['scenes']} x {args.adjoint_n} spheres, {adjoint['steps']} steps: forward "
              f"{adjoint['forward_s']:.2f}s, forward+checkpoints {adjoint['forward_with_checkpoints_s']:.2f}s, "
              f"reverse {adjoint['backward_s']:.2f}s, ratio {adjoint['adjoint_over_forward']:.2f} "
              f"({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        from dem_mbd_cuda_amd import _abi, host

        lib = _abi.Library(os.path.join(ROOT, "oracle", "liboracle.so"), "ora_")
        o = lib.solver()
        o.initialize(sc)
        c0 = time.perf_counter()
        o.compute_forces(collect=True)  # the initial evaluation of the reference loop (untimed)
        o.init_half_step(0.5 * sc.dt)
        c1 = time.perf_counter()
        for _ in range(args.cpu_steps):
            o.advance(sc.dt)
            o.compute_forces(collect=False)
            o.finish(sc.dt)
        c2 = time.perf_counter()
        o.close()
        cpu = {"value": sc.n * args.cpu_steps / (c2 - c1), "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"CPU oracle (C restatement of the reference law, 1 thread) on the same "
                         f"{sc.n}-sphere {args.packing} scene: {args.cpu_steps} full step(s) timed "
                         f"({c2 - c1:.1f} s) after the untimed initial evaluation ({c1 - c0:.1f} s)"}
        print(f"# cpu oracle: {cpu['value']:.3g} particle-steps/s ({c2 - c1:.1f} s timed, {c1 - c0:.1f} s set-up)",
              file=sys.stderr, flush=True)

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
