"""Multi-step runs (dem_run_steps): steps replayed from captured HIP graphs,
with and without the autotuner, must equal the step-by-step entry points
bit for bit.  The graphs are replayed many times with the program's own
downloads and synchronisations in between, and every chunk is compared."""
import numpy as np
import pytest

from dem_mbd_cuda_amd import scene

pytestmark = pytest.mark.gpu


def box(n=800, seed=9):
    sc = scene.box_scene(n=n, seed=seed, box=(0.04, 0.04, 0.04), steps=200)
    rng = np.random.RandomState(seed)
    sc.vel = rng.uniform(-0.4, 0.4, (sc.n, 3))
    sc.angvel = rng.uniform(-40, 40, (sc.n, 3))
    return sc


def snapshot(s):
    st = s.state()
    fo = s.forces()
    h = s.debug_history()
    return [st["pos"], st["vel"], st["angvel"], st["quat"], fo["contact"], fo["wall"], fo["total"],
            h["pp_keys"], h["pp_values"], h["pw_keys"], h["pw_values"]]


def start(lib, sc):
    s = lib.solver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * sc.dt)
    return s


@pytest.mark.parametrize("skin", ["0.2", "0"])
@pytest.mark.parametrize("reorder", ["0", "1", "3"])
def test_graph_replays_equal_eager_steps(hip, reorder, skin, monkeypatch):
    """Multi-step runs (candidate lists with skin 0.2, or the full scan from
    captured graphs with skin 0), whose steps finish and advance their rows in
    the law kernel while another step follows, against step-by-step calls."""
    monkeypatch.setenv("DEM_REORDER_INTERVAL", reorder)
    monkeypatch.setenv("DEM_SKIN", skin)
    sc = box()
    a, b = start(hip, sc), start(hip, sc)
    dt = sc.dt
    for chunk in (1, 17, 40, 40, 23):
        assert a.run_steps(chunk, dt, graph=True) == chunk
        for _ in range(chunk):
            b.advance(dt)
            b.compute_forces(collect=False)
            b.finish(dt)
        for x, y in zip(snapshot(a), snapshot(b)):
            assert np.array_equal(x, y)
    assert a.stats()["particle_contacts"] == b.stats()["particle_contacts"]
    a.close()
    b.close()


def test_autotuned_run_equals_eager_steps(hip):
    sc = box(n=1200, seed=3)
    a, b = start(hip, sc), start(hip, sc)
    dt = sc.dt
    assert a.run_steps(120, dt, graph=True, autotune=True) == 120
    for _ in range(120):
        b.advance(dt)
        b.compute_forces(collect=False)
        b.finish(dt)
    for x, y in zip(snapshot(a), snapshot(b)):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("skin", ["0.2", "0"])
def test_batch_graph_equals_eager(hip, skin, monkeypatch):
    """A batch of scenes with their own walls (the law kernel's fused pass
    reads each row's scene walls) through multi-step runs, against
    step-by-step calls."""
    from test_gpu_batch import agitated_sweep

    monkeypatch.setenv("DEM_SKIN", skin)
    scenes = agitated_sweep(count=4, seed=7, nmin=200, nmax=900)
    batch = scene.SceneBatch(scenes)
    a, b = hip.solver(), hip.solver()
    for s in (a, b):
        s.initialize_batch(batch)
        s.compute_forces(collect=True)
        s.init_half_step(0.5 * batch.combined.dt)
    dt = batch.combined.dt
    for chunk in (30, 20):
        assert a.run_steps(chunk, dt) == chunk
        for _ in range(chunk):
            b.advance(dt)
            b.compute_forces(collect=False)
            b.finish(dt)
        for x, y in zip(snapshot(a), snapshot(b)):
            assert np.array_equal(x, y)
    assert a.stats()["particle_contacts"] > 0
    assert np.abs(a.forces()["wall"]).max() > 0.0
    a.close()
    b.close()
