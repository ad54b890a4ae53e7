"""The benchmark configurations themselves on the HIP path, against the
oracle and (for the batch) against single-scene runs:

* the 1M-sphere hopper (lognormal radii clipped to [0.5, 2] mm, floor and
  four inclined walls, cell 2 x max radius) in its settled bed (sequential
  deposition): the initial evaluation and two steps bit for bit -- cell keys
  and the device's own sorted order of the initial evaluation, contact-pair
  sets, history tables, states, forces;
* a 30,000-sphere settled hopper bed with the same radii, walls and cell, agitated, over
  100 steps with a small candidate-list skin, so that lists are rebuilt
  several times (each rebuild sorts on the device), bit for bit;
* all 256 scenes of the batched parameter sweep (N ~ U[500, 20000], per-scene
  mu, restitution, E), agitated so that contacts form, 40 steps in one
  batch: every scene equals its own single-scene run bit for bit, and a
  seeded sample of 8 scenes equals the oracle bit for bit.
"""
import numpy as np
import pytest

from helpers import run_frames
from test_gpu_parity import compare_frames

pytestmark = pytest.mark.gpu


def test_hopper_1m_initial_evaluation_and_two_steps(hip, oracle):
    from dem_mbd_cuda_amd import scene

    sc = scene.hopper_scene(n=1_000_000, packing="settled", mesh_scale=2.0, steps=2)
    print("1M hopper built", flush=True)
    g = run_frames(hip, sc, steps=2, interval=2, debug=True)
    print("HIP run done", flush=True)
    o = run_frames(oracle, sc, steps=2, interval=2, debug=True)
    print("oracle run done", flush=True)
    assert g[0]["counters"]["last_sorted"] == 1  # frame 0 compares the device's own sort
    assert o[2]["stats"]["particle_contacts"] > 3 * sc.n
    compare_frames(g, o)


def test_hopper_lists_rebuilt_on_the_headline_cell(hip, oracle, monkeypatch):
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_SKIN", "0.0005")  # rebuild threshold 0.5 um
    sc = scene.hopper_scene(n=30_000, width=0.2 * 0.03 ** 0.5, packing="settled", mesh_scale=2.0, steps=100)
    rng = np.random.RandomState(41)
    sc.vel = rng.uniform(-0.05, 0.05, (sc.n, 3))
    sc.angvel = rng.uniform(-50.0, 50.0, (sc.n, 3))
    g = run_frames(hip, sc, steps=100, interval=25, debug=True)
    o = run_frames(oracle, sc, steps=100, interval=25, debug=True)
    c = g[100]["counters"]
    print("counters", c, flush=True)
    assert c["rebuilds"] >= 3 and c["sorts"] >= 3, c
    assert o[100]["stats"]["particle_contacts"] > 3 * sc.n
    compare_frames(g, o)


def _agitate(scenes, seed):
    rng = np.random.RandomState(seed)
    for sc in scenes:
        sc.vel = rng.uniform(-3.0, 3.0, (sc.n, 3))
        sc.angvel = rng.uniform(-100.0, 100.0, (sc.n, 3))
    return scenes


def _run(lib, sc, steps, batch=None):
    from dem_mbd_cuda_amd import host

    s = lib.solver()
    if batch is not None:
        s.initialize_batch(batch)
        sc = batch.combined
    else:
        s.initialize(sc)
    host.run_pure(s, sc, steps=steps)
    out = dict(s.state(), **{f"f_{k}": v for k, v in s.forces().items() if v is not None and k != "field"})
    out["contacts"] = s.stats()["particle_contacts"]
    s.close()
    return out


def test_sweep_256_scenes_batch_equals_single_runs_and_oracle(hip, oracle):
    from dem_mbd_cuda_amd import scene

    steps = 40
    scenes = _agitate(scene.sweep_scenes(count=256, seed=4), 43)
    batch = scene.SceneBatch(scenes)
    print("sweep built:", batch.n, "spheres", flush=True)
    b = _run(hip, None, steps, batch=batch)
    assert b["contacts"] > 0.05 * batch.n
    cut = lambda a, k: a[batch.particle_start[k]:batch.particle_start[k + 1]]
    keys = [k for k in b if k != "contacts"]
    for k, sc in enumerate(scenes):
        one = _run(hip, sc, steps)
        for name in keys:
            assert np.array_equal(cut(b[name], k), one[name]), (k, name)
    print("batch == single runs for all 256 scenes", flush=True)
    sample = np.random.RandomState(44).choice(len(scenes), 8, replace=False)
    for k in sample:
        ref = _run(oracle, scenes[k], steps)
        assert ref["contacts"] > 0, k
        for name in keys:
            assert np.array_equal(cut(b[name], k), ref[name]), (k, name)
    print("batch == oracle on scenes", sorted(int(k) for k in sample), flush=True)
