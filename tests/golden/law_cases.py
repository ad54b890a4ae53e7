"""Case directories for the contact-law and integrator pins against the
reference's legacy CPU binary (tests/test_law_pins.py).

Every case runs the reference's multi-body loop (CONTROL_MULTIBODY_FLAG = 1),
whose particle integration is the CUDA backend's scheme: the half-step
quaternion q' = N(q + gamma(q, w) dt/2) and q += gamma(q', w_half) dt
(reference src/simulation.cpp IntegrateDemMultiBody, legacy branch,
against src/cuda/cuda_particle_backend.cu:149-265), with the same Euler
equations (src/dyneq.cpp:5-27 against cuda_particle_backend.cu:71-101).
Gravity is off and no contact exists at the first evaluation, so the CUDA
law's zero-length first evaluation (cuda_backend_cuda.cu:402-409) never
meets a contact; spheres keep constant velocities until their first contact,
where the CUDA law's half-step and end-point rates coincide.

    write_case(name, directory)   writes <directory>/InputFile/...
    CASES                         name -> case writer
"""
from __future__ import annotations

import os

import numpy as np

MECH = dict(mu=0.4, mu_T=0.3, mu_R=0.5, epsN=0.6, epsS=0.7, cohesive=0, beta=0.6, kN=2.0e4)


def _fmt(v):
    return "%.17g" % float(v)


def _quat(seed):
    q = np.random.RandomState(seed).normal(size=4)
    return q / np.linalg.norm(q)


def _rot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _write(path, text):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


def _particles(d, rows):
    """rows: (id, status, mass, radius, pos3, vel3, quat4, omega_body3)."""
    lines = []
    for pid, st, m, r, p, v, q, w in rows:
        vals = [m, r, *p, *v, *q, *w]
        lines.append(f"{pid} {st} " + " ".join(_fmt(x) for x in vals))
    _write(os.path.join(d, "InputFile", "Particles", "Particles.bt"), "\n".join(lines) + "\n")


def _inertia(d, rows):
    _write(os.path.join(d, "InputFile", "Particles", "ParticleInertia.bt"),
           "".join(f"{pid} {_fmt(a)} {_fmt(b)} {_fmt(c)}\n" for pid, (a, b, c) in rows))


def _walls(d, walls):
    text = f"{len(walls)}\n" + "".join(" ".join(_fmt(x) for x in (*o, *n)) + "\n" for o, n in walls)
    _write(os.path.join(d, "InputFile", "Environment", "Wall.bt"), text)


def _body(d, verts, faces, mass, center, vel, omega, state):
    """One rigid body of triangles (vertices in the body frame, 1-based
    faces), identity orientation and identity inertia."""
    mesh = f"{len(verts)} {len(faces)}\n" + "".join(" ".join(_fmt(x) for x in v) + "\n" for v in verts)
    mesh += "".join(" ".join(str(i) for i in f) + "\n" for f in faces)
    _write(os.path.join(d, "InputFile", "BodySet", "body.bt"), mesh)
    head = "filepath,mass,x,y,z,vx,vy,vz,omega_x,omega_y,omega_z,lambda_x,lambda_y,lambda_z,lambda_w," \
           "I1,I2,I3,I4,I5,I6,I7,I8,I9,state"
    row = ["Data\\DATA\\InputFile\\BodySet\\body.bt", mass, *center, *vel, *omega, 1, 0, 0, 0,
           1, 0, 0, 0, 1, 0, 0, 0, 1, state]
    _write(os.path.join(d, "InputFile", "BodySet", "bodies.csv"),
           "1" + "," * 24 + "\n" + head + "\n" + ",".join(_fmt(x) if not isinstance(x, str) else x
                                                          for x in row[:-1]) + f",{state}\n")


def _par(d, dt, steps, interval, walls=False, inertia=False):
    par = {"ODE_Option": 1, "ODE_StepSize": _fmt(dt), "ODE_EndStep": steps, "ODE_StartStep": 0,
           "ODE_OutputInterval": interval}
    par.update({f"MECH_{k}": (_fmt(v) if isinstance(v, float) else v) for k, v in MECH.items()})
    par.update({"CONTROL_GRAVITY_FLAG": 0, "CONTROL_GRAVITY_X": 0, "CONTROL_GRAVITY_Y": 0, "CONTROL_GRAVITY_Z": 0,
                "CONTROL_WALL_FLAG": 1 if walls else 0,
                "CONTROL_WALL_FILE": "Data\\DATA\\InputFile\\Environment\\Wall.bt",
                "CONTROL_MULTIBODY_FLAG": 1,
                "CONTROL_MULTIBODY_FILE": "Data\\DATA\\InputFile\\BodySet\\bodies.csv",
                "CONTROL_UNIVERSAL_GRAVITATION_FLAG": 0,
                "CONTROL_PARTICLES_FILE": "Data\\DATA\\InputFile\\Particles\\Particles.bt"})
    if inertia:
        par["CONTROL_PARTICLE_INERTIA_FILE"] = "Data\\DATA\\InputFile\\Particles\\ParticleInertia.bt"
    par.update({"CONTROL_OUTPUT_PARTICLE_STATE": 1, "CONTROL_OUTPUT_PARTICLE_FORCE": 1,
                "CONTROL_OUTPUT_BODY_STATE": 1, "CONTROL_OUTPUT_BODY_FORCE": 1, "CONTROL_OUTPUT_SUMMARY": 0})
    _write(os.path.join(d, "InputFile", "SettingData", "world.par"),
           "".join(f"{k} = {v}\n" for k, v in par.items()))


def _far_plate(d):
    """A fixed one-triangle body 1 m away (the multi-body loop needs a body)."""
    _body(d, [(0, 0, 0), (0.01, 0, 0), (0, 0.01, 0)], [(1, 2, 3)], 1.0, (1.0, 1.0, 1.0), (0, 0, 0), (0, 0, 0), 1)


def free_spinners(d):
    """Six free particles far apart: asymmetric principal inertia (sidecar),
    seeded orientations and body-frame spins of 20-120 rad/s, 2,000 steps of
    1e-5 s, frames every 250 steps.  No force acts: the spins follow Euler's
    torque-free equations and the orientations their quaternion kinematics."""
    rng = np.random.RandomState(11)
    rows, inert = [], []
    for k in range(6):
        m, r = 2.0e-3, 1.0e-3
        base = 0.4 * m * r * r
        I = base * np.array([1.0, 1.0 + 0.3 * (k + 1), 0.5 + 0.4 * k])
        w = rng.uniform(-120.0, 120.0, 3)
        rows.append((k, 0, m, r, (0.05 * k, 0.0, 0.0), (0.01 * k, -0.02, 0.005), _quat(100 + k), w))
        inert.append((k, I))
    _particles(d, rows)
    _inertia(d, inert)
    _far_plate(d)
    _par(d, 1.0e-5, 2000, 250, inertia=True)


def wall_slider(d):
    """Two spheres of asymmetric inertia and seeded orientation sliding and
    spinning onto an inclined plane (a parallel plane 0.1 m above it stays
    clear), first touching at the fourth evaluation, then 86 steps in
    contact, frames every step.  The particle-wall law takes end-point rates
    only, as the legacy law does, so every frame compares."""
    nrm = np.array([0.3, -0.2, 1.0])
    nrm /= np.linalg.norm(nrm)
    walls = [((0.0, 0.0, 0.0), tuple(nrm)), (tuple(0.1 * nrm), tuple(-nrm))]
    t1 = np.cross(nrm, [0.0, 1.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(nrm, t1)
    rows, inert = [], []
    for k in range(2):
        m, r = 4.0e-3, 1.2e-3
        vn, vt = 0.08, 0.05
        gap = 3.5 * vn * 1e-6
        pos = nrm * (r + gap) + (0.01 * k) * t2 + 1e-3 * t1
        vel = -vn * nrm + vt * (t1 if k == 0 else t2)
        q = _quat(200 + k)
        w_world = np.array([3.0, -25.0, 8.0]) if k == 0 else np.array([-12.0, 6.0, 20.0])
        w_body = _rot(q).T @ w_world
        base = 0.4 * m * r * r
        rows.append((k, 0, m, r, pos, vel, q, w_body))
        inert.append((k, base * np.array([1.0, 1.4, 0.8])))
    _particles(d, rows)
    _inertia(d, inert)
    _walls(d, walls)
    _far_plate(d)
    _par(d, 1.0e-6, 90, 1, walls=True, inertia=True)


def unequal_pair(d):
    """Two spheres of radii 1.5 mm and 0.9 mm (solid-sphere inertia) closing
    obliquely with spins about different axes, touching at the fifth
    evaluation; frames every step up to it.  The legacy law takes end-point
    rates throughout and the CUDA law half-step rates for the springs, which
    are equal until the first contact acts, so that frame compares."""
    m1, r1, m2, r2 = 3.0e-3, 1.5e-3, 1.2e-3, 0.9e-3
    dirv = np.array([1.0, 0.35, -0.2])
    dirv /= np.linalg.norm(dirv)
    vrel = -0.1 * dirv + np.array([0.0, 0.04, 0.03])
    closing = -float(vrel @ dirv)
    gap = 4.5 * closing * 1e-6
    p2 = dirv * (r1 + r2 + gap)
    rows = [(0, 0, m1, r1, (0, 0, 0), (0, 0, 0), _quat(300), (40.0, -15.0, 25.0)),
            (1, 0, m2, r2, p2, vrel, _quat(301), (-30.0, 50.0, 10.0))]
    _particles(d, rows)
    _far_plate(d)
    _par(d, 1.0e-6, 5, 1)


def body_slider(d):
    """A sphere sliding and spinning onto one large triangle of a prescribed
    body that translates and turns (omega_z = 3 rad/s) about a centre 30 mm
    away from the contact; first touching after a few steps, frames every
    step up to the first contact evaluation (the fourth).  The body-particle detail (FB),
    the body loads (At) and the sphere's force and couple compare there."""
    # triangle in the body frame around the future contact, facing +z
    c = np.array([-0.03, 0.01, 0.0])
    verts = [c + (-0.02, -0.02, 0.0), c + (0.03, -0.015, 0.0), c + (-0.01, 0.03, 0.0)]
    center = (0.03, -0.01, 0.0)
    m, r = 4.0e-3, 1.2e-3
    vn = 0.08
    gap = 3.5 * vn * 1e-6
    rows = [(0, 0, m, r, (0.0, 0.0, r + gap), (0.03, -0.02, -vn), _quat(400), (20.0, -35.0, 15.0))]
    _particles(d, rows)
    _body(d, verts, [(1, 2, 3)], 50.0, center, (0.0, 0.0, 0.0), (0.0, 0.0, 3.0), 2)
    _par(d, 1.0e-6, 4, 1)


CASES = {
    "free_spinners": free_spinners,
    "wall_slider": wall_slider,
    "unequal_pair": unequal_pair,
    "body_slider": body_slider,
}


def write_case(name, directory):
    CASES[name](directory)
    return directory
