"""Clump components (owners made of 1 to 8 sphere components): the sidecar
readers against the reference's schema test (tests/particle_schema_test.cpp),
and the oracle's clump semantics pinned by invariants that a plausible slip
(owner instead of component centre, a missing offset x force torque, a
wrong transport of the angular velocity, contacts between components of one
owner, a component summed twice) would break:

  * a one-component owner with a body-frame offset feels, at the first
    evaluation, the force of a sphere placed at the component centre and
    moving with the component's velocity v + w x r; its torque is that
    sphere's torque plus r x F;
  * a clump feels the sum of the forces its components would feel as such
    spheres, pair by pair, and components of one owner never touch.
"""
import os

import numpy as np
import pytest

from dem_mbd_cuda_amd import scene


# ---- sidecar readers (particle_schema_test.cpp:49-113) -------------------
def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


PIDS = np.array([31, 47], np.int32)


def test_principal_inertia_sidecar(tmp_path):
    p = _write(tmp_path / "ParticleInertia.bt", "31 0.01 0.02 0.03\n\n47 0.04 0.05 0.06\n")
    inertia = scene.read_principal_inertia(p, PIDS)
    assert inertia[0, 1] == 0.02 and inertia[1, 2] == 0.06


@pytest.mark.parametrize("text,message", [
    ("31 0.01 0.02 0.03\n47 0.04 -0.05 0.06\n", "finite and strictly positive"),
    ("31 0.01 0.02 0.03\n", "exactly one record per owner"),
    ("31 0.01 0.02 0.03\n31 0.01 0.02 0.03\n", "Duplicate owner ID"),
    ("31 0.01 0.02 0.03\n48 0.04 0.05 0.06\n", "Unknown owner ID"),
    ("31 0.01 0.02\n47 0.04 0.05 0.06\n", "expected owner_id Ixx Iyy Izz"),
    ("31 0.01 0.02 0.03 9\n47 0.04 0.05 0.06\n", "expected owner_id Ixx Iyy Izz"),
])
def test_principal_inertia_rejects(tmp_path, text, message):
    p = _write(tmp_path / "bad_inertia.bt", text)
    with pytest.raises(RuntimeError, match=message):
        scene.read_principal_inertia(p, PIDS)


def test_clump_component_sidecar(tmp_path):
    p = _write(tmp_path / "ClumpComponents.bt", "101 31 0.08 -0.02 0 0\n102 31 0.08 0.02 0 0\n201 47 0.20 0 0 0\n")
    c = scene.read_clump_components(p, PIDS)
    assert len(c["comp_id"]) == 3
    assert c["comp_owner"][1] == 0 and c["comp_id"][2] == 201
    assert c["comp_pos_body"][0, 0] == -0.02 and c["comp_radius"][2] == 0.20


@pytest.mark.parametrize("text,message", [
    ("101 31 0.08 0 0 0\n101 47 0.20 0 0 0\n", "Duplicate component ID"),
    ("-1 31 0.08 0 0 0\n201 47 0.20 0 0 0\n", r"Clump component IDs must lie in \[0, INT32_MAX\]"),
    ("".join(f"{100 + k} 31 0.01 0 0 0\n" for k in range(9)) + "201 47 0.20 0 0 0\n", "at most 8"),
    ("101 31 0.08 0 0 0\n", "missing owner 47"),
    ("101 31 0.0 0 0 0\n201 47 0.20 0 0 0\n", "finite with positive radius"),
    ("101 31 0.08 0 nan 0\n201 47 0.20 0 0 0\n", "finite with positive radius"),
    ("101 99 0.08 0 0 0\n201 47 0.20 0 0 0\n", "Unknown owner ID"),
    ("101 31 0.08 0 0\n201 47 0.20 0 0 0\n", "expected component_id owner_id"),
    ("\n \n", "must not be empty"),
])
def test_clump_component_rejects(tmp_path, text, message):
    p = _write(tmp_path / "bad_components.bt", text)
    with pytest.raises(RuntimeError, match=message):
        scene.read_clump_components(p, PIDS)


def test_case_reader_loads_sidecars(tmp_path):
    """world.par names both sidecars (simulation.cpp:243-246, 109-110)."""
    d = tmp_path / "case"
    data = d / "InputFile"
    os.makedirs(data / "SettingData")
    _write(data / "Particles.bt", "31 0 2.0 0.1 0 0 0 0 0 0 1 0 0 0 0 0 0\n"
                                  "47 0 3.0 0.2 1 0 0 0 0 0 1 0 0 0 0 0 0\n")
    _write(data / "Inertia.bt", "31 0.01 0.02 0.03\n47 0.04 0.05 0.06\n")
    _write(data / "Clumps.bt", "101 31 0.08 -0.02 0 0\n102 31 0.08 0.02 0 0\n201 47 0.20 0 0 0\n")
    _write(data / "SettingData" / "world.par",
           "ODE_StepSize = 1e-6\nODE_EndStep = 10\nCONTROL_PARTICLES_FILE = InputFile/Particles.bt\n"
           "CONTROL_PARTICLE_INERTIA_FILE = InputFile/Inertia.bt\n"
           "CONTROL_CLUMP_COMPONENTS_FILE = InputFile/Clumps.bt\n")
    sc = scene.read_case(str(d))
    assert sc.inertia[0, 0] == 0.01 and sc.inertia[1, 2] == 0.06
    assert list(sc.comp_id) == [101, 102, 201] and list(sc.comp_owner) == [0, 0, 1]
    # without the sidecars: solid-sphere inertia 0.4 m r^2 and one component per owner
    _write(data / "SettingData" / "world.par",
           "ODE_StepSize = 1e-6\nODE_EndStep = 10\nCONTROL_PARTICLES_FILE = InputFile/Particles.bt\n")
    sc = scene.read_case(str(d))
    assert abs(sc.inertia[0, 0] - 0.008) <= 1e-15
    assert list(sc.comp_id) == [31, 47] and list(sc.comp_owner) == [0, 1]


# ---- oracle clump semantics ----------------------------------------------
def _quat_rotate(q, v):
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, float)


MECH = scene.mech(epsS=0.6, mu=0.4, epsN=0.7, cohesion=0.0, beta=0.5, muT=0.2, muR=0.1, kN=2.0e4)


def _first_forces(oracle, sc):
    s = oracle.solver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    f = s.forces()
    st = s.stats()
    s.close()
    return f, st


def _owner(pid, pos, vel, angvel, quat, mass=1.0e-3):
    return dict(pid=pid, pos=pos, vel=vel, angvel=angvel, quat=quat, mass=mass)


def _build(owners, comps, walls=None):
    """owners: list of dicts; comps: list of (id, owner row, radius, body offset)."""
    n = len(owners)
    P = dict(pid=np.array([o["pid"] for o in owners], np.int32), status=np.zeros(n, np.int32),
             mass=np.array([o["mass"] for o in owners]), radius=np.array([max(c[2] for c in comps)] * n),
             inertia=np.full((n, 3), 2.0e-9), pos=np.array([o["pos"] for o in owners], float),
             vel=np.array([o["vel"] for o in owners], float), angvel=np.array([o["angvel"] for o in owners], float),
             quat=np.array([o["quat"] for o in owners], float))
    kw = {}
    if walls is not None:
        kw = dict(wall_origin=np.array([walls[0]], float), wall_normal=np.array([walls[1]], float))
    return scene.Scene(**P, comp_id=np.array([c[0] for c in comps], np.int32),
                       comp_owner=np.array([c[1] for c in comps], np.int32),
                       comp_radius=np.array([c[2] for c in comps], float),
                       comp_pos_body=np.array([c[3] for c in comps], float).reshape(-1, 3),
                       pp=MECH, pt=dict(MECH), pw=dict(MECH), dt=1e-5, **kw)


def _as_sphere(owner, comp):
    """The sphere equivalent of one component: at the component centre, with
    the component's velocity, the owner's angular velocity, mass and pose."""
    q = np.asarray(owner["quat"], float)
    off = _quat_rotate(q, comp[3])
    w_world = _quat_rotate(q, owner["angvel"])
    o = dict(owner)
    o["pos"] = np.asarray(owner["pos"], float) + off
    o["vel"] = np.asarray(owner["vel"], float) + np.cross(w_world, off)
    return o, (comp[0], 0, comp[2], (0.0, 0.0, 0.0)), off


Q1 = tuple(np.array([0.9, 0.2, -0.3, 0.25]) / np.linalg.norm([0.9, 0.2, -0.3, 0.25]))
Q2 = tuple(np.array([0.5, -0.6, 0.4, 0.2]) / np.linalg.norm([0.5, -0.6, 0.4, 0.2]))


def test_offset_component_equals_displaced_sphere(oracle):
    r = 1.0e-3
    A = _owner(3, (0.0, 0.0, 0.0), (0.02, -0.01, 0.005), (40.0, -25.0, 10.0), Q1)
    compA = (30, 0, r, (0.4e-3, -0.3e-3, 0.2e-3))
    offA = _quat_rotate(Q1, compA[3])
    B = _owner(5, tuple(np.asarray(A["pos"]) + offA + np.array([1.9e-3, 0.3e-3, -0.2e-3])), (-0.01, 0.0, 0.02),
               (-5.0, 15.0, 30.0), Q2, mass=1.5e-3)
    compB = (50, 1, r, (0.0, 0.0, 0.0))
    wall = ((0.0, 0.0, offA[2] - 0.9e-3), (0.0, 0.0, 1.0))  # touches A's component only
    f, st = _first_forces(oracle, _build([A, B], [compA, compB], wall))
    assert st["particle_contacts"] == 2
    As, compAs, off = _as_sphere(A, compA)
    g, _ = _first_forces(oracle, _build([As, B], [compAs, compB], wall))
    for k in ("contact", "wall"):
        F = g[k][0, :3]
        assert np.allclose(f[k][0, :3], F, rtol=1e-12, atol=0.0), k
        assert np.allclose(f[k][0, 3:], g[k][0, 3:] + np.cross(off, F), rtol=1e-9, atol=1e-18), k
        assert np.abs(F).max() > 0.0, k
    assert np.allclose(f["contact"][1], g["contact"][1], rtol=1e-12, atol=0.0)


def test_clump_force_is_the_sum_over_component_pairs(oracle):
    r = 1.0e-3
    A = _owner(3, (0.0, 0.0, 0.0), (0.01, 0.0, -0.02), (10.0, 20.0, -30.0), Q1)
    B = _owner(4, (1.3e-3, 0.2e-3, 0.1e-3), (-0.02, 0.01, 0.0), (-20.0, 5.0, 15.0), Q2, mass=2.0e-3)
    # two overlapping components each (they overlap their own sibling)
    cA = [(30, 0, r, (-0.5e-3, 0.0, 0.0)), (31, 0, 0.9 * r, (0.6e-3, 0.1e-3, 0.0))]
    cB = [(40, 1, r, (0.4e-3, 0.0, 0.2e-3)), (41, 1, 0.8 * r, (-0.5e-3, 0.2e-3, 0.0))]
    f, st = _first_forces(oracle, _build([A, B], cA + cB))
    # every component pair as a two-sphere scene
    F = np.zeros((2, 6))
    pairs = 0
    for a in cA:
        for b in cB:
            As, ca, offa = _as_sphere(A, a)
            Bs, cb, offb = _as_sphere(B, b)
            cb = (cb[0], 1, cb[2], cb[3])
            g, gst = _first_forces(oracle, _build([As, Bs], [ca, cb]))
            if gst["particle_contacts"] == 0:
                continue
            pairs += 1
            for row, off in ((0, offa), (1, offb)):
                F[row, :3] += g["contact"][row, :3]
                F[row, 3:] += g["contact"][row, 3:] + np.cross(off, g["contact"][row, :3])
    assert pairs >= 2
    assert st["particle_contacts"] == 2 * pairs  # both owners count each pair; siblings never
    assert np.allclose(f["contact"], F, rtol=1e-10, atol=1e-18)


def test_siblings_never_touch(oracle):
    """Three heavily overlapping components of one owner and nothing else:
    no contact, no force."""
    r = 1.0e-3
    A = _owner(1, (0.0, 0.0, 0.0), (0.1, 0.0, 0.0), (0.0, 50.0, 0.0), Q2)
    comps = [(10, 0, r, (0.0, 0.0, 0.0)), (11, 0, r, (0.3e-3, 0.0, 0.0)), (12, 0, r, (0.0, 0.3e-3, 0.0))]
    f, st = _first_forces(oracle, _build([A], comps))
    assert st["particle_contacts"] == 0
    assert not np.any(f["contact"])


def test_clumpify_keeps_components_inside_owner():
    sc = scene.box_scene(n=50, seed=3, box=(0.02, 0.02, 0.02), steps=1)
    cl = scene.clumpify(sc, np.arange(50) % 8 + 1, seed=2)
    assert len(cl.comp_id) == sum(np.arange(50) % 8 + 1)
    reach = np.linalg.norm(cl.comp_pos_body, axis=1) + cl.comp_radius
    assert np.all(reach <= sc.radius[cl.comp_owner] * (1 + 1e-12))
    assert len(set(cl.comp_id.tolist())) == len(cl.comp_id)
