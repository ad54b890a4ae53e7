"""Batched scenes (dem_initialize_batch): every scene of a batch must come
out bit for bit as when the scene runs alone in its own solver -- states,
force outputs and contact counts -- for scenes of different sizes, walls and
laws whose row ranges straddle thread blocks."""
import numpy as np
import pytest

from dem_mbd_cuda_amd import host, scene

pytestmark = pytest.mark.gpu


def agitated_sweep(count, seed, nmin, nmax):
    """Sweep scenes with radii inflated 5 % after insertion (initial overlaps
    and wall contacts) and random translational / angular velocities."""
    rng = np.random.RandomState(seed)
    out = []
    for sc in scene.sweep_scenes(count=count, seed=seed, nmin=nmin, nmax=nmax, dt=2e-6):
        n = sc.n
        r = sc.radius * 1.05
        p = scene.sphere_particles(sc.pos, r, density=2500.0, ids=sc.pid)
        p["vel"] = rng.uniform(-1.0, 1.0, (n, 3))
        p["angvel"] = rng.uniform(-100.0, 100.0, (n, 3))
        out.append(scene.Scene(**p, wall_origin=sc.wall_origin, wall_normal=sc.wall_normal,
                               pp=sc.pp, pt=sc.pt, pw=sc.pw, dt=sc.dt, gravity=sc.gravity,
                               gravity_enabled=1, steps=120))
    return out


def run_batch(lib, batch, steps):
    s = lib.solver()
    s.initialize_batch(batch)
    host.run_pure(s, batch.combined, steps=steps)
    st, fo, stats = s.state(), s.forces(), s.stats()
    s.close()
    return st, fo, stats


def run_single(lib, sc, steps):
    s = lib.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=steps)
    st, fo, stats = s.state(), s.forces(), s.stats()
    s.close()
    return st, fo, stats


def test_batch_equals_single_scenes(hip):
    scenes = agitated_sweep(count=5, seed=21, nmin=300, nmax=2500)
    batch = scene.SceneBatch(scenes)
    steps = 120
    bst, bfo, bstats = run_batch(hip, batch, steps)
    contacts = 0
    for k, sc in enumerate(scenes):
        st, fo, stats = run_single(hip, sc, steps)
        contacts += stats["particle_contacts"]
        for key in ("pos", "vel", "angvel", "quat"):
            assert np.array_equal(batch.split(bst[key])[k], st[key]), (k, key)
        for key in ("contact", "wall", "field", "total"):
            assert np.array_equal(batch.split(bfo[key])[k], fo[key]), (k, key)
    assert contacts > 0
    assert bstats["particle_contacts"] == contacts


def test_batch_of_one_equals_single(hip):
    sc = agitated_sweep(count=1, seed=5, nmin=800, nmax=800)[0]
    bst, bfo, _ = run_batch(hip, scene.SceneBatch([sc]), 60)
    st, fo, _ = run_single(hip, sc, 60)
    for key in ("pos", "vel", "angvel", "quat"):
        assert np.array_equal(bst[key], st[key]), key
    for key in ("contact", "wall", "total"):
        assert np.array_equal(bfo[key], fo[key]), key
