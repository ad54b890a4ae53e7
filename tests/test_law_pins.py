"""Pins of the oracle's contact laws and integrator that do not come from the
oracle (CPU only).

1. Against the reference's own legacy CPU binary (fixtures written by
   tests/golden/make_fixtures.py from oracle/_ref/dem_mbd_cpu on the cases of
   tests/golden/law_cases.py).  The reference's multi-body loop integrates
   particles with the CUDA backend's scheme (half-step quaternion, Euler
   equations) and evaluates the legacy law, which equals the CUDA law where
   the cases stay: a contact's first evaluation with constant velocities
   before it (particle-particle, particle-triangle), and particle-wall
   contacts throughout (end-point rates in both).  Every number of every
   frame file the case runner writes must equal the binary's to 1e-10 of its
   vector's size (the files carry 12 to 17 digits):
     free_spinners  Euler's equations (cuda_particle_backend.cu:71-101), the
                    quaternion derivative and the half-step quaternion
                    (:149-265) over 2,000 steps of six free asymmetric bodies;
     wall_slider    the particle-wall law with its rolling rate
                    (cuda_contact_backend.cu:975-1063), the torque's turn into
                    the body frame, over 86 steps of sliding, spinning contact;
     unequal_pair   the particle-particle law at first contact, unequal radii:
                    the couple li n x Ft (:641-699), twist and roll;
     body_slider    the triangle manifold and the per-triangle contact at
                    first contact (:708-949): the sphere's force and couple,
                    the body-particle detail, the body loads with the friction
                    torque -(arm x F) about an off-centre body turning at
                    3 rad/s.
2. Closed forms the mathematics fixes: a sphere spinning about a fixed axis
   turns by q(t) = (cos(wt/2), sin(wt/2) axis); the first step of a free
   asymmetric body changes its spin by dt times Euler's torque-free
   w' = -I^-1 (w x I w); and a free asymmetric body keeps its world angular
   momentum R(q) I w and its kinetic energy to the integrator's second order.

tests/test_oracle_mutants.py plants single-line mistakes in copies of the
oracle and requires each to fail one of these pins.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN
from dem_mbd_cuda_amd import host, run, scene

import sys

sys.path.insert(0, GOLDEN)
import law_cases  # noqa: E402

REL = 1e-10


def _groups(kind, ncols):
    """Column groups (vectors) of the reference's frame files."""
    if kind == "ph":
        return [[0], [1], [2], [3], [4, 5, 6], [7, 8, 9], [10, 11, 12, 13], [14, 15, 16]]
    if kind == "At":
        return [[0], [1, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12, 13], [14, 15, 16], [17, 18, 19]]
    if kind == "FB":
        return [[3 * k, 3 * k + 1, 3 * k + 2] for k in range(6)] + [[18]]
    return [[3 * k, 3 * k + 1, 3 * k + 2] for k in range(ncols // 3)]


def _frame_files(case_dir):
    base = os.path.join(case_dir, "OutputFile")
    out = {}
    for sub in sorted(os.listdir(base)):
        if not os.path.isdir(os.path.join(base, sub)):
            continue
        for f in sorted(os.listdir(os.path.join(base, sub))):
            if f.endswith(".bt"):
                out[f"{sub}/{f}"] = np.loadtxt(os.path.join(base, sub, f), ndmin=2)
    return out


def legacy_pin(lib, name, tmp_path):
    """Runs law case `name` through the case runner on `lib` and compares
    every frame file with the reference binary's numbers."""
    ref = np.load(os.path.join(GOLDEN, "law_pins_ref_cpu.npz"))
    d = str(tmp_path / f"{name}_{abs(hash(lib.path)) % 100000}")
    shutil.rmtree(d, ignore_errors=True)
    law_cases.write_case(name, d)
    run.run_case(d, lib=lib, log=lambda *a: None)
    mine = _frame_files(d)
    keys = sorted(k.split("/", 1)[1] for k in ref.files if k.startswith(name + "/"))
    assert keys and sorted(mine) == keys, (name, sorted(mine)[:4], keys[:4])
    worst = (0.0, "")
    for k in keys:
        b = ref[f"{name}/{k}"]
        a = mine[k][:, :b.shape[1]]  # the legacy At file has no impulse columns
        assert a.shape == b.shape, (k, a.shape, b.shape)
        kind = os.path.basename(k).split(".")[0]
        for g in _groups(kind, b.shape[1]):
            scale = np.maximum(np.max(np.abs(b[:, g]), axis=1), 1e-300)
            err = np.max(np.abs(a[:, g] - b[:, g]), axis=1) / scale
            if err.max() > worst[0]:
                worst = (float(err.max()), f"{k} columns {g}")
    assert worst[0] <= REL, f"{name}: {worst[1]} differs from the reference CPU binary by {worst[0]:.3e}"
    return worst[0]


@pytest.mark.parametrize("name", list(law_cases.CASES))
def test_legacy_binary_pins(oracle, name, tmp_path):
    legacy_pin(oracle, name, tmp_path)


def test_law_cases_reach_their_contacts(oracle, tmp_path):
    """The compared frames hold the contacts the pins are about: wall
    contact from the fourth evaluation on, the pair and the triangle contact
    at the last frame, with rolling and twisting moments present."""
    ref = np.load(os.path.join(GOLDEN, "law_pins_ref_cpu.npz"))
    w = ref["wall_slider/force_particles/F.00000004.bt"]
    assert np.all(np.linalg.norm(w[:, 6:9], axis=1) > 0) and np.all(ref["wall_slider/force_particles/F.00000003.bt"][:, 6:12] == 0)
    p = ref["unequal_pair/force_particles/F.00000005.bt"]
    assert np.all(np.linalg.norm(p[:, 0:3], axis=1) > 0) and np.all(ref["unequal_pair/force_particles/F.00000004.bt"][:, 0:6] == 0)
    fb = ref["body_slider/force_bodys/FB.00004.bt"]
    assert np.linalg.norm(fb[0, 12:15]) > 0 and np.linalg.norm(fb[0, 15:18]) > 0 and fb[0, 18] > 0
    assert np.all(ref["body_slider/force_bodys/FB.00003.bt"] == 0)


# ---- closed forms ---------------------------------------------------------
def _free_scene(inertia, omega_body, quat, dt, steps):
    P = scene.sphere_particles([[0.0, 0.0, 0.0]], [1e-3], density=2500.0)
    P["inertia"] = np.asarray(inertia, float).reshape(1, 3)
    P["angvel"] = np.asarray(omega_body, float).reshape(1, 3)
    P["quat"] = np.asarray(quat, float).reshape(1, 4)
    return scene.Scene(**P, dt=dt, steps=steps)


def _rot(q):
    return law_cases._rot(q)


def spin_closed_form(lib):
    """A sphere spinning at 100 rad/s about a fixed unit axis, 1,000 steps of
    1e-5 s (one radian): q(t) = (cos(wt/2), sin(wt/2) axis) to 1e-7 (the
    step's own error is about 3e-9 here)."""
    axis = np.array([0.48, -0.6, 0.64])
    w, dt, steps = 100.0, 1e-5, 1000
    sc = _free_scene([4e-10] * 3, w * axis, [1, 0, 0, 0], dt, steps)
    s = lib.solver()
    s.initialize(sc)
    host.run_pure(s, sc)
    q = s.state()["quat"][0]
    s.close()
    t = dt * steps
    want = np.concatenate([[np.cos(w * t / 2)], np.sin(w * t / 2) * axis])
    assert np.max(np.abs(q - want)) < 1e-7, (q, want)


def euler_first_step(lib):
    """The first step of a free body with principal moments (1, 2, 3) x 1e-9
    and a turned orientation: (w1 - w0) / dt = -I^-1 (w0 x I w0)."""
    I = np.array([1.0, 2.0, 3.0]) * 1e-9
    w0 = np.array([30.0, -20.0, 50.0])
    dt = 1e-6
    sc = _free_scene(I, w0, law_cases._quat(7), dt, 1)
    s = lib.solver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    s.init_half_step(0.5 * dt)
    s.advance(dt)
    w1 = s.state()["angvel"][0]
    s.close()
    alpha = -np.cross(w0, I * w0) / I
    got = (w1 - w0) / dt
    assert np.max(np.abs(got - alpha)) <= 1e-7 * np.max(np.abs(alpha)), (got, alpha)


def free_body_invariants(lib, dt=1e-5):
    """A free asymmetric body, 2,000 steps of 1e-5 s (about 3 turns): the
    world angular momentum R(q) I w and the kinetic energy stay within 2e-6
    (relative) of their initial values (about 2e-7 and 4e-7 here); halving
    the step cuts the drift of the momentum four times (second order)."""
    I = np.array([1.0, 1.7, 2.6]) * 1e-9
    w0 = np.array([60.0, -90.0, 40.0])
    q0 = law_cases._quat(9)
    drifts = []
    for h in (dt, dt / 2):
        steps = int(round(0.02 / h))
        sc = _free_scene(I, w0, q0, h, steps)
        s = lib.solver()
        s.initialize(sc)
        host.run_pure(s, sc)
        st = s.state()
        s.close()
        q, w = st["quat"][0], st["angvel"][0]
        L0, L = _rot(q0) @ (I * w0), _rot(q) @ (I * w)
        E0, E = 0.5 * np.sum(I * w0 * w0), 0.5 * np.sum(I * w * w)
        drifts.append((np.linalg.norm(L - L0) / np.linalg.norm(L0), abs(E - E0) / E0))
    (dl1, de1), (dl2, de2) = drifts
    assert dl1 < 2e-6 and de1 < 2e-6, drifts
    assert 3.5 < dl1 / max(dl2, 1e-30) < 4.5, drifts


def test_spin_about_a_fixed_axis(oracle):
    spin_closed_form(oracle)


def test_euler_equations_at_the_first_step(oracle):
    euler_first_step(oracle)


def test_free_body_invariants_second_order(oracle):
    free_body_invariants(oracle)
