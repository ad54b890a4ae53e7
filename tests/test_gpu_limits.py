"""Capacity limits and range errors on the HIP path, against the oracle: a
sphere with more contacts than history slots (the reference drops the
contacts that find no slot and reports the overflow on the next diagnostics
step), more listed partners than a candidate list holds, more touching walls
than wall slots, and a component outside the 21-bit cell range.  States,
forces and history tables stay bit-exact where the run continues; errors
carry the reference's messages."""
import numpy as np
import pytest

from dem_mbd_cuda_amd import scene
from dem_mbd_cuda_amd._abi import DemError

pytestmark = pytest.mark.gpu

MECH = scene.mech(epsS=0.5, mu=0.4, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=2.0e4)


def crowded(n_small, gap, seed=3):
    """One 10 mm sphere and n_small 1 mm spheres around it, each `gap` from
    touching it (negative = overlapping), on seeded random directions."""
    rng = np.random.RandomState(seed)
    d = rng.normal(size=(n_small, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    R, r = 10e-3, 1e-3
    pos = np.vstack([[0.0, 0.0, 0.0], d * (R + r + gap)])
    radius = np.concatenate([[R], np.full(n_small, r)])
    P = scene.sphere_particles(pos, radius, density=2500.0)
    P["vel"] = rng.uniform(-0.05, 0.05, (n_small + 1, 3))
    P["angvel"] = rng.uniform(-20, 20, (n_small + 1, 3))
    return scene.Scene(**P, pp=MECH, pt=dict(MECH), pw=dict(MECH), dt=1e-6, steps=0)


def run_steps(lib, sc, steps):
    """Initial evaluation and `steps` steps without diagnostics, then one
    diagnostics evaluation; returns the frame and the error text (or None)."""
    s = lib.solver()
    s.initialize(sc)
    s.compute_forces(collect=False)
    s.init_half_step(0.5 * sc.dt)
    for _ in range(steps):
        s.advance(sc.dt)
        s.compute_forces(collect=False)
        s.finish(sc.dt)
    err = None
    s.advance(sc.dt)
    try:
        s.compute_forces(collect=True)
    except DemError as e:
        err = str(e)
    frame = dict(state=s.state(), forces=s.forces(), history=s.debug_history(), grid=s.debug_grid())
    s.close()
    return frame, err


def compare(g, o):
    for k in ("pos", "vel", "angvel", "quat"):
        assert np.array_equal(g["state"][k], o["state"][k]), k
    for k in ("contact", "wall", "total"):
        assert np.array_equal(g["forces"][k], o["forces"][k]), k
    for k in ("pp_keys", "pp_values", "pw_keys", "pw_values"):
        assert np.array_equal(g["history"][k], o["history"][k]), k
    assert np.array_equal(g["grid"][0], o["grid"][0]) and np.array_equal(g["grid"][1], o["grid"][1])


@pytest.mark.parametrize("skin", ["0.2", "0"])
def test_more_contacts_than_history_slots(hip, oracle, skin, monkeypatch):
    """80 spheres overlapping one: 16 of its contacts find no slot; with and
    without candidate lists (DEM_SKIN=0 forces the full scan)."""
    monkeypatch.setenv("DEM_SKIN", skin)
    sc = crowded(80, -0.05e-3)
    g, ge = run_steps(hip, sc, 10)
    o, oe = run_steps(oracle, sc, 10)
    assert oe is not None and "particle contact-history capacity exceeded" in oe
    assert ge == oe
    compare(g, o)


def test_more_partners_than_a_candidate_list(hip, oracle):
    """100 spheres overlapping one: the first rebuild overflows the 96-entry
    candidate list and the solver goes on with the full scan, where the
    contacts also outnumber the history slots; results and the error stay
    the oracle's."""
    sc = crowded(100, -0.05e-3, seed=5)
    g, ge = run_steps(hip, sc, 20)
    o, oe = run_steps(oracle, sc, 20)
    assert oe is not None and "particle contact-history capacity exceeded" in oe
    assert ge == oe
    compare(g, o)


def test_more_touching_walls_than_wall_slots(hip, oracle):
    """18 planes through a sphere: the 17th and 18th find no wall slot."""
    rng = np.random.RandomState(7)
    nrm = rng.normal(size=(18, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    P = scene.sphere_particles([[0.0, 0.0, 0.0]], [1e-3], density=2500.0)
    sc = scene.Scene(**P, wall_origin=-nrm * 0.9e-3, wall_normal=nrm, pp=MECH, pt=dict(MECH), pw=dict(MECH),
                     dt=1e-6, steps=0)
    g, ge = run_steps(hip, sc, 3)
    o, oe = run_steps(oracle, sc, 3)
    assert oe is not None and "wall/body contact-history capacity exceeded" in oe
    assert ge == oe
    compare(g, o)


def test_cell_range_error(hip, oracle):
    """A sphere 1e6 m out: its cell index leaves the signed 21-bit range."""
    P = scene.sphere_particles([[0.0, 0.0, 0.0], [1.0e6, 0.0, 0.0]], [1e-3, 1e-3], density=2500.0)
    sc = scene.Scene(**P, pp=MECH, pt=dict(MECH), pw=dict(MECH), dt=1e-6, steps=0)
    msgs = []
    for lib in (hip, oracle):
        s = lib.solver()
        s.initialize(sc)
        with pytest.raises(DemError) as e:
            s.compute_forces(collect=True)
        msgs.append(str(e.value))
        s.close()
    assert msgs[0] == msgs[1] and "21-bit" in msgs[0]


def _runaway():
    """A sphere 2,000 m out moving at 1e8 m/s (100 m a step): with 2.5 mm
    cells it leaves the signed 21-bit range (2,621 m) on the seventh step."""
    P = scene.sphere_particles([[0.0, 0.0, 0.0], [2000.0, 0.0, 0.0]], [1e-3, 1e-3], density=2500.0)
    P["vel"] = np.array([[0.0, 0.0, 0.0], [1.0e8, 0.0, 0.0]])
    return scene.Scene(**P, pp=MECH, pt=dict(MECH), pw=dict(MECH), dt=1e-6, steps=0)


def _step_until_error(s, sc, steps, collect_last):
    """Initial evaluation and `steps` steps without diagnostics (the last one
    with diagnostics if asked); returns (step index of the error, text)."""
    s.initialize(sc)
    s.compute_forces(collect=False)
    s.init_half_step(0.5 * sc.dt)
    for k in range(1, steps + 1):
        s.advance(sc.dt)
        try:
            s.compute_forces(collect=collect_last and k == steps)
        except DemError as e:
            return k, str(e)
        s.finish(sc.dt)
    return None, None


@pytest.mark.parametrize("skin", ["0.2", "0"])
def test_cell_range_error_mid_run(hip, oracle, monkeypatch, skin):
    """A sphere leaves the 21-bit range on a step without diagnostics.  The
    oracle (like the reference) fails at that step.  With candidate lists the
    count comes back with the displacement flag, so the HIP path fails at the
    same step; on full-scan steps it fails once the step's counters reach the
    host, at the latest on the next diagnostics step."""
    monkeypatch.setenv("DEM_SKIN", skin)
    sc = _runaway()
    so = oracle.solver()
    ko, eo = _step_until_error(so, sc, 12, collect_last=True)
    so.close()
    assert ko == 7 and "21-bit" in eo
    s = hip.solver()
    kg, eg = _step_until_error(s, sc, 12, collect_last=True)
    assert eg == eo
    if skin != "0":
        assert kg == ko
    else:
        assert ko <= kg <= 12
    s.close()


def test_cell_range_error_in_graph_run(hip, oracle, monkeypatch):
    """The same runaway sphere in a multi-step run of graph replays (full
    scan, no bodies): the run fails at its end with the reference's message,
    having run its steps; the oracle fails at the seventh step."""
    monkeypatch.setenv("DEM_SKIN", "0")
    sc = _runaway()
    msgs = []
    for lib in (hip, oracle):
        s = lib.solver()
        s.initialize(sc)
        s.compute_forces(collect=False)
        s.init_half_step(0.5 * sc.dt)
        with pytest.raises(DemError) as e:
            s.run_steps(40, sc.dt, graph=True)
        msgs.append(str(e.value))
        if lib is hip:
            c = s.run_counters()
            assert c["graph_steps"] > 0, c
        s.close()
    assert msgs[0] == msgs[1] and "21-bit" in msgs[0]


def test_stage_timings_only_when_enabled(hip):
    """Stage events are recorded only while enabled: the last evaluation made
    with timings on reports five non-negative stage times, and nothing is
    reported once they are turned off."""
    sc = crowded(40, -2e-5)
    s = hip.solver()
    s.initialize(sc)
    s.compute_forces(collect=False)
    assert len(s.last_timings()) == 0
    s.enable_timings(True)
    s.compute_forces(collect=False)
    t = s.last_timings()
    assert len(t) == 5 and np.all(t >= 0.0) and t.sum() > 0.0
    s.enable_timings(False)
    s.compute_forces(collect=False)
    assert len(s.last_timings()) == 0
    s.close()
