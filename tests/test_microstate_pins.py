"""Closed-form pins for the 22 body-microstate sums
(observeBodyMicrostateKernel, cuda_contact_backend.cu:1226-1356) on the
oracle, from the reference's own triangle validation case (a sphere of
radius 0.1 pressed 0.01 into one triangle with kN = 1e4: |Fn| = 100 N, its
section disc wholly inside the triangle, so the area weight is 1):

  * at the first evaluation (no spring yet, contact centre at the body
    centre) every field has a closed form;
  * sliding and spinning on a body whose centre is off the contact, the
    fields follow from the stored tangential and twisting springs (read from
    the body history table), the contact force and the lever arm."""
import math

import numpy as np

from conftest import golden
from dem_mbd_cuda_amd import host

MU, MUT, BETA, R = 0.3, 0.1, 0.5, 0.1   # the case's MECH block and radius


def micro_after(oracle, sc, steps):
    s = oracle.solver()
    s.initialize(sc)
    out = {}

    def cb(step, t, extra):
        if step + 1 == steps:
            out["micro"] = s.stats()["micro_instant"][0].copy()
            out["force"] = s.forces()["total"][0].copy()  # triangle force only (no gravity, walls, pairs)
            out["hist"] = s.debug_history()
            out["state"] = s.state()
            out["bodies"] = extra["bodies"]

    host.run_coupled(s, sc, steps=steps, output_interval=1, on_frame=cb)
    s.close()
    return out


def test_first_evaluation_closed_form(oracle):
    sc = golden("case_triangle_contact.npz")
    sc.body_state = np.array([1], np.int32)
    o = micro_after(oracle, sc, 0)
    v = o["micro"]
    fn = 100.0
    expect = np.zeros(22)
    expect[0] = fn                       # |Fn| * weight
    expect[1] = fn * R                   # |Fn| r
    expect[7] = MU * MUT * BETA * R * fn  # twisting limit
    expect[9] = MU * fn                  # sliding limit
    expect[12] = 1.0                     # area weight
    expect[15] = -fn                     # load on the body, -Fn (z)
    assert np.allclose(v, expect, rtol=1e-12, atol=1e-12), (v, expect)


def test_sliding_on_an_off_centre_body(oracle):
    sc = golden("case_triangle_contact.npz")
    sc.body_state = np.array([1], np.int32)  # fixed body
    c = np.array([0.3, 0.2, 0.0])            # body centre off the contact; same world triangle
    pose = np.asarray(sc.body_pose, float).reshape(-1, 13).copy()
    pose[0, 0:3] = c
    sc.body_pose = pose
    tl = np.asarray(sc.tri_local, float).reshape(-1, 15).copy()
    for v in range(4):
        tl[:, 3 * v:3 * v + 3] -= c
    sc.tri_local = tl
    sc.vel = np.array([[0.02, -0.01, 0.0]])
    sc.angvel = np.array([[0.0, 0.0, 3.0]])
    steps = 4
    o = micro_after(oracle, sc, steps)
    v = o["micro"]
    keys, vals = o["hist"]["pt_keys"][0], o["hist"]["pt_values"][0]
    slot = int(np.nonzero(keys == 1)[0][0])
    th, wh = vals[slot, 3:6], vals[slot, 6:9]
    assert np.linalg.norm(th) > 0.0 and np.linalg.norm(wh) > 0.0
    p = o["state"]["pos"][0]
    cc = np.array([p[0], p[1], 0.0])                  # contact centre on the plane z = 0
    arm = cc - c
    radial = arm - np.array([0.0, 0.0, 1.0]) * arm[2]
    fnn = v[0]                                         # |Fn|, weight 1
    assert abs(v[12] - 1.0) <= 1e-14
    assert abs(fnn - o["force"][2]) <= 1e-9 * fnn      # the normal force of the step
    tlim, wlim = MU * fnn, MU * MUT * BETA * R * fnn
    np.testing.assert_allclose(v[1], fnn * R, rtol=1e-13)
    np.testing.assert_allclose(v[2], fnn * np.linalg.norm(radial), rtol=1e-12)
    np.testing.assert_allclose(v[3:6], -wh, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(v[6], np.linalg.norm(wh), rtol=1e-12)
    np.testing.assert_allclose(v[7], wlim, rtol=1e-13)
    np.testing.assert_allclose(v[8], np.linalg.norm(th), rtol=1e-12)
    np.testing.assert_allclose(v[9], tlim, rtol=1e-13)
    np.testing.assert_allclose(v[10], min(1.0, np.linalg.norm(wh) / wlim), rtol=1e-12)
    np.testing.assert_allclose(v[11], min(1.0, np.linalg.norm(th) / tlim), rtol=1e-12)
    np.testing.assert_allclose(v[13:16], [0.0, 0.0, -fnn], rtol=1e-12, atol=1e-9 * fnn)
    np.testing.assert_allclose(v[16:19], -th, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(v[19:22], np.cross(arm, -th), rtol=1e-10, atol=1e-18)
