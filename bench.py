"""FP64 DEM step throughput on one MI355X (particle-steps/s).

Workload (default): pure DEM, 1,000,000 spheres, radius lognormal (median
1 mm, sigma 0.2, clipped to [0.5, 2] mm), density 2500 kg/m^3, kN = E * median
radius with E = 1e7 Pa, mu 0.5, restitution 0.5, seeded random sequential
insertion into a 200 x 200 mm hopper (floor + 4 walls inclined 10 degrees),
gravity -9.81 z, dt = 5e-7 s, grid cell = 2 x max radius.  --packing dense
replaces the loose insertion with a near-touching bed (contact-heavy).
A step is the reference's pure-DEM loop body: advanceParticles,
computeForces, finishParticleStep, all state resident in HBM.

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (the contact-law kernel, timed with HIP events on the solver's stream)
and a CPU baseline (the CPU oracle, single thread, on a bounded sample of the
same workload).
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


def law_bytes_per_particle(cbar: float, walls: int) -> float:
    """Algorithmic HBM bytes of the contact-law kernel per particle: own state
    read (24 f64 + 7 i32 = 220 B), contact list count (4 B), history masks
    (PP/PW occupancy + seen words read and written, 48 B), outputs (27 f64 =
    216 B), wall parameters are cached; per stored contact 4 B list entry,
    4 B key and 72 B history values read + 72 B written (152 B)."""
    return 220.0 + 4.0 + 48.0 + 216.0 + 152.0 * cbar


def run_steps(s, dt, steps):
    for _ in range(steps):
        s.advance(dt)
        s.compute_forces(collect=False)
        s.finish(dt)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=60,
                    help="untimed steps; the first ones autotune the reorder interval and contact grid")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mesh-scale", type=float, default=2.0,
                    help="grid cell / max radius (the workload specifies 2; the reference uses 2.5)")
    ap.add_argument("--packing", choices=("rsa", "dense"), default="rsa",
                    help="rsa: the specified seeded insertion (loose); dense: near-touching bed")
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

    import numpy as np

    import dem_mbd_cuda_amd as dm
    from dem_mbd_cuda_amd import scene

    t0 = time.time()
    sc = scene.hopper_scene(n=args.n, seed=2, mesh_scale=args.mesh_scale, packing=args.packing)
    print(f"# scene built in {time.time() - t0:.1f}s, {sc.n} spheres", file=sys.stderr, flush=True)
    dt = sc.dt
    s = dm.CudaDemSolver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    # steps run as dem_run_steps: (advance, computeForces, finish) replayed from
    # captured HIP graphs; the warmup autotunes the reorder interval and grid
    s.run_steps(args.warmup, dt, graph=True, autotune=True)
    s.synchronize()
    print(f"# warmup done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    s.run_steps(args.steps, dt, graph=True)
    s.synchronize()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    value = sc.n * args.steps / elapsed
    # kernel timing (HIP events on the solver's stream) over a few more steps
    tms = []
    for _ in range(5):
        s.advance(dt)
        s.compute_forces(collect=False)
        s.finish(dt)
        tms.append(s.last_timings())
    tm = np.mean(np.array(tms), axis=0)
    grid, scan, pairs, law = tm[0], tm[1], tm[2], tm[3]
    s.advance(dt)
    s.compute_forces(collect=True)  # a diagnostics step for the contact count
    s.finish(dt)
    st = s.stats()
    cbar = st["particle_contacts"] / sc.n
    law_ms = float(pairs + law)
    bytes_law = law_bytes_per_particle(cbar, len(sc.wall_origin)) * sc.n
    achieved = bytes_law / (law_ms * 1e-3) / 1e9
    print(f"# ms/step {1e3 * elapsed / args.steps:.3f}; grid {grid:.3f} ms, scan {scan:.3f} ms, "
          f"pairs {pairs:.3f} ms, sums {law:.3f} ms; contacts/particle {cbar:.3f}",
          file=sys.stderr, flush=True)
    s.close()

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
        host.run_pure(o, sc, steps=args.cpu_steps)
        c1 = time.perf_counter()
        o.close()
        # run_pure does one extra initial force evaluation
        cpu = {"value": sc.n * args.cpu_steps / (c1 - c0), "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"CPU oracle (C restatement of the reference law, 1 thread) on the same "
                         f"{sc.n}-sphere scene, initial force evaluation + {args.cpu_steps} steps"}

    out = {
        "metric": "FP64 particle-steps/s @1M",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded RSA packing)",
        "config": {"workload": f"pure DEM hopper, {sc.n} spheres lognormal r (median 1 mm, sigma 0.2, "
                               f"clipped 0.5-2 mm), floor + 4 walls inclined 10 deg, gravity, dt 5e-7, "
                               f"grid cell {args.mesh_scale} x max radius, "
                               + ("seeded random sequential insertion" if args.packing == "rsa"
                                  else "dense near-touching packing"),
                   "particles": sc.n, "dt": dt, "packing": args.packing,
                   "contacts_per_particle": cbar},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_law", "kernel_ms": law_ms,
                     "bytes_per_particle": law_bytes_per_particle(cbar, len(sc.wall_origin))},
        "cpu_baseline": cpu,
        "batched_sweep": sweep,
        "adjoint": adjoint,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
