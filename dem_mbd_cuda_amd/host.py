"""CPU step driver around the solver: the reference's pure-DEM and DEM-MBD
time loops (src/cuda_simulation_integrator.cpp:245-446) and the rigid-body
Newton-Euler evaluation (src/dyneq.cpp:97-129), with the same call order,
half-step bookkeeping and frame numbering (initial frame index 0, the frame
after step s has index s + 1).

Frames are handed to a callback as dictionaries instead of being written to
files, so the same loop drives tests, benchmarks and file writers.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

BODY_DYNAMIC, BODY_FIXED, BODY_PRESCRIBED = 0, 1, 2


def gama(q, w):
    """Quaternion derivative 0.5 * q (x) (0, w) (reference matrixtrans.cpp:87-96)."""
    return np.array([0.5 * (-q[1] * w[0] - q[2] * w[1] - q[3] * w[2]),
                     0.5 * (q[0] * w[0] + q[2] * w[2] - q[3] * w[1]),
                     0.5 * (q[0] * w[1] - q[1] * w[2] + q[3] * w[0]),
                     0.5 * (q[0] * w[2] + q[1] * w[1] - q[2] * w[0])])


def qnormalize(q):
    """vector4d::normalize: unit quaternion, zero if |q| <= 1e-10 (matrixtrans.h:208-216)."""
    length = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    if length > 1e-10:
        return q * (1.0 / length)
    return np.zeros(4)


def quat2dcm(q):
    """Quat2DCM (reference matrixtrans.cpp:36-63)."""
    qq = [q[0] * q[0], q[1] * q[1], q[2] * q[2], q[3] * q[3], 2.0 * q[0] * q[1],
          2.0 * q[0] * q[2], 2.0 * q[0] * q[3], 2.0 * q[1] * q[2], 2.0 * q[1] * q[3],
          2.0 * q[2] * q[3]]
    return np.array([[qq[0] + qq[1] - qq[2] - qq[3], qq[7] + qq[6], qq[8] - qq[5]],
                     [qq[7] - qq[6], qq[0] - qq[1] + qq[2] - qq[3], qq[9] + qq[4]],
                     [qq[8] + qq[5], qq[9] - qq[4], qq[0] - qq[1] - qq[2] + qq[3]]])


def matvec(m, v):
    return np.array([m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2],
                     m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
                     m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]])


def dyn_eqn_bodyset(bodies, F):
    """DynEqnBodySet (reference dyneq.cpp:97-129): [b][6] accelerations."""
    out = np.zeros((len(bodies["state"]), 6))
    for i, st in enumerate(bodies["state"]):
        if st != BODY_DYNAMIC:
            continue
        m = bodies["mass"][i]
        out[i, :3] = [F[i][j] / m for j in range(3)]
        T = matvec(quat2dcm(bodies["orien"][i]), F[i][3:6])
        I = bodies["inertia"][i]
        w = bodies["angvel"][i]
        out[i, 3] = (T[0] - (I[2][2] - I[1][1]) * w[2] * w[1]) / I[0][0]
        out[i, 4] = (T[1] - (I[0][0] - I[2][2]) * w[0] * w[2]) / I[1][1]
        out[i, 5] = (T[2] - (I[1][1] - I[0][0]) * w[1] * w[0]) / I[2][2]
    return out


def _pose13(b):
    return np.concatenate([b["center"], b["orien"], b["vel"], b["angvel"]], axis=1)


def restored(solver):
    """hasRestoredParticleLeapfrogState: a DMHIST2 checkpoint was loaded."""
    f = getattr(solver, "restored_leapfrog", None)
    return bool(f()) if f is not None else False


def run_pure(solver, scene, steps=None, output_interval=None, on_frame=None, start_step=0,
             history_in=None, history_out=None):
    """IntegrateDem (reference cuda_simulation_integrator.cpp:245-286).
    history_in: contact-history checkpoint loaded before the first force
    evaluation when the file exists; history_out: written after the last
    step (the reference's InputFile/ and OutputFile/contact_history.bt)."""
    if history_in is not None and os.path.exists(history_in):
        solver.load_history(history_in)
    dt = scene.dt
    half = 0.5 * dt
    steps = scene.steps if steps is None else steps
    interval = output_interval or scene.output_interval or steps
    end = start_step + steps
    solver.compute_forces(collect=True)
    if on_frame:
        on_frame(-1, 0.0, None)
    if not restored(solver):
        solver.init_half_step(half)
    counter = 0
    for step in range(start_step, end):
        next_time = (step + 1) * dt
        solver.advance(dt)
        due = counter + 1 == interval or step == end - 1
        solver.compute_forces(collect=due)
        solver.finish(dt, half)
        counter += 1
        if counter == interval or step == end - 1:
            if on_frame:
                on_frame(step, next_time, None)
            counter = 0
    if history_out is not None:
        solver.save_history(history_out)


def load_prescribed_motion(path, start_step, end_step):
    """loadPrescribedMotion (reference cuda_simulation_integrator.cpp:180-221):
    rows "step vx vy vz wx wy wz" ('#' comments, commas as separators; rows
    whose first field is not an integer are skipped); steps strictly
    increasing inside [start_step, end_step), the first one at start_step.
    Returns [(step, linear[3], angular[3])] ([] for an empty path)."""
    if not path:
        return []
    try:
        f = open(path, encoding="utf-8", errors="replace")
    except OSError:
        raise RuntimeError(f"cannot open prescribed motion file: {path}")
    keys = []
    with f:
        for number, line in enumerate(f, start=1):
            line = line.split("#", 1)[0].replace(",", " ")
            tokens = line.split()
            if not tokens:
                continue
            m = re.match(r"[+-]?\d+", tokens[0])
            if not m or not -2 ** 31 <= int(m.group(0)) < 2 ** 31:
                continue  # not an int: the reference's `row >> key.step` fails and skips the row
            rest = tokens[0][m.end():]
            tokens = ([rest] if rest else []) + tokens[1:]
            try:
                vals = [float(t) for t in tokens[:6]]
            except ValueError:
                vals = []
            if len(vals) < 6:
                raise RuntimeError(f"invalid prescribed motion row {number}")
            if len(tokens) > 6:
                raise RuntimeError(f"extra prescribed motion column at row {number}")
            step = int(m.group(0))
            if (keys and step <= keys[-1][0]) or step < start_step or step >= end_step:
                raise RuntimeError("prescribed motion steps must increase within the integration range")
            keys.append((step, np.array(vals[:3]), np.array(vals[3:])))
    if not keys or keys[0][0] != start_step:
        raise RuntimeError("prescribed motion must begin at ODE_StartStep")
    return keys


def check_prescribed_bodies(scene):
    """The one-prescribed-body rule of applyPrescribedMotion (reference
    cuda_simulation_integrator.cpp:229-232), for a scene about to run a
    schedule."""
    if len(scene.body_mass) != 1 or int(scene.body_state[0]) != BODY_PRESCRIBED:
        raise RuntimeError("prescribed motion schedule currently requires one prescribed body")


def _apply_prescribed(keys, cursor, step, B):
    """applyPrescribedMotion (reference cuda_simulation_integrator.cpp:223-243)
    without the half-step part: advances the cursor to the last key at or
    before `step` and sets the one prescribed body's velocities from it.
    Returns the cursor."""
    while cursor + 1 < len(keys) and keys[cursor + 1][0] <= step:
        cursor += 1
    if len(B["mass"]) != 1 or B["state"][0] != BODY_PRESCRIBED:
        raise RuntimeError("prescribed motion schedule currently requires one prescribed body")
    B["vel"][0] = keys[cursor][1]
    B["angvel"][0] = keys[cursor][2]
    return cursor


def body_state_from_scene(scene):
    nb = len(scene.body_mass)
    pose = np.asarray(scene.body_pose, float).reshape(nb, 13)
    return dict(center=pose[:, 0:3].copy(), orien=pose[:, 3:7].copy(), vel=pose[:, 7:10].copy(),
                angvel=pose[:, 10:13].copy(), mass=np.asarray(scene.body_mass, float),
                inertia=np.asarray(scene.body_inertia, float).reshape(nb, 3, 3),
                state=np.asarray(scene.body_state, int))


def run_coupled(solver, scene, steps=None, output_interval=None, on_frame=None, start_step=0,
                history_in=None, history_out=None, prescribed=None):
    """IntegrateDemMultiBody (reference cuda_simulation_integrator.cpp:288-446).
    on_frame(step, time, extra): extra holds the body state, the loads and
    the impulses of that frame.  history_in / history_out as for run_pure;
    prescribed: a load_prescribed_motion schedule (applyPrescribedMotion,
    :223-243, at the start of every step)."""
    if history_in is not None and os.path.exists(history_in):
        solver.load_history(history_in)
    dt = scene.dt
    half = 0.5 * dt
    steps = scene.steps if steps is None else steps
    interval = output_interval or scene.output_interval or steps
    end = start_step + steps
    B = body_state_from_scene(scene)
    nb = len(B["mass"])
    impulses = np.zeros((nb, 6))
    cursor = 0
    if prescribed:
        # the schedule's first key sets the body before the first evaluation (:292-294)
        cursor = _apply_prescribed(prescribed, cursor, start_step, B)
    loads = solver.compute_forces(_pose13(B), loads=True, collect=True)
    if on_frame:
        on_frame(-1, 0.0, dict(bodies={k: v.copy() for k, v in B.items()}, loads=loads.copy(),
                               impulses=impulses.copy()))
    if not restored(solver):
        solver.init_half_step(half)
    acc = dyn_eqn_bodyset(B, loads)
    hv = np.zeros((nb, 3)); hw = np.zeros((nb, 3)); hq = np.zeros((nb, 4)); hdq = np.zeros((nb, 4))
    bacc = np.zeros((nb, 3)); bdom = np.zeros((nb, 3))

    def half_update(first):
        for b in range(nb):
            st = B["state"][b]
            bacc[b] = acc[b, :3]
            bdom[b] = acc[b, 3:]
            if st == BODY_DYNAMIC:
                if first:
                    hv[b] = B["vel"][b] + bacc[b] * half
                    hw[b] = B["angvel"][b] + bdom[b] * half
                else:
                    hv[b] = hv[b] + bacc[b] * dt
                    hw[b] = hw[b] + bdom[b] * dt
            elif st == BODY_FIXED:
                if first:
                    bacc[b] = 0.0
                    bdom[b] = 0.0
                else:
                    B["vel"][b] = 0.0
                    B["angvel"][b] = 0.0
                hv[b] = 0.0
                hw[b] = 0.0
            else:
                if first:
                    bacc[b] = 0.0
                    bdom[b] = 0.0
                hv[b] = B["vel"][b]
                hw[b] = B["angvel"][b]
            hq[b] = qnormalize(B["orien"][b] + gama(B["orien"][b], B["angvel"][b]) * half)
            hdq[b] = gama(hq[b], hw[b])

    half_update(True)
    counter = 0
    advanced = False  # the step's particle advance was launched at the end of the previous one
    for step in range(start_step, end):
        if prescribed:
            cursor = _apply_prescribed(prescribed, cursor, step, B)
            hv[0] = B["vel"][0]
            hw[0] = B["angvel"][0]
            hq[0] = qnormalize(B["orien"][0] + gama(B["orien"][0], B["angvel"][0]) * half)
            hdq[0] = gama(hq[0], hw[0])
        current = step * dt
        next_time = (step + 1) * dt
        if not advanced:
            solver.advance(dt)
        advanced = False
        for b in range(nb):
            if B["state"][b] == BODY_FIXED:
                continue
            B["center"][b] = B["center"][b] + hv[b] * dt
            B["orien"][b] = qnormalize(B["orien"][b] + hdq[b] * dt)
            B["vel"][b] = hv[b] + bacc[b] * half
            B["angvel"][b] = hw[b] + bdom[b] * half
        due = counter + 1 == interval or step == end - 1
        loads = solver.compute_forces(_pose13(B), loads=True, collect=due)
        impulses += loads * dt
        solver.finish(dt, half)
        if not due and step + 1 < end:
            # the particle finish and the next advance need no body state:
            # launched now, they run on the device during the body solve below
            solver.advance(dt)
            advanced = True
        acc = dyn_eqn_bodyset(B, loads)
        half_update(False)
        counter += 1
        if counter == interval or step == end - 1:
            if on_frame:
                on_frame(step, current, dict(bodies={k: v.copy() for k, v in B.items()},
                                             loads=loads.copy(), impulses=impulses.copy()))
            impulses[:] = 0.0
            counter = 0
    if history_out is not None:
        solver.save_history(history_out)
    return B
