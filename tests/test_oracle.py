"""Pins for the CPU oracle (CPU only).

The oracle restates the reference's CUDA law; these checks tie it to things
that do not come from the oracle itself:
  - the closed forms of the reference's scenario_validator on its own
    validation inputs (tests/scenario_validator.cpp:75-113);
  - the sphere/triangle area known answers of tests/contact_geometry_test.cpp;
  - the component transform values of tests/component_transform_test.cu:66-102;
  - the reference's own CPU binary (oracle/_ref/dem_mbd_cpu) on DATA1 at zero
    velocity, where the legacy law and the CUDA law coincide;
  - mechanics the law must satisfy whatever its implementation: linear
    spring-slider accumulation up to the Coulomb cap, the restitution
    coefficient of the linear dashpot, action-reaction and angular momentum
    balance of pair forces, cell-key ordering.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden
from helpers import assert_close, run_frames

from dem_mbd_cuda_amd import host, scene


def frames_of(oracle, case, interval=1):
    sc = golden(f"case_{case}.npz")
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    return sc, run_frames(oracle, sc, interval=interval)


# ---- scenario_validator closed forms (tests/scenario_validator.cpp:75-113) --
def test_free_fall(oracle):
    sc, f = frames_of(oracle, "free_fall")
    st = f[10]["state"]
    assert abs(st["pos"][0, 2] - (1.0 + 0.5 * -9.8 * 0.01 * 0.01)) <= 2e-12
    assert abs(st["vel"][0, 2] - (-9.8 * 0.01)) <= 2e-12
    assert abs(f[0]["forces"]["field"][0, 2] - (-9.8)) <= 2e-12


def test_wall_contact(oracle):
    sc, f = frames_of(oracle, "wall_contact")
    assert abs(f[0]["forces"]["wall"][0, 2] - 100.0) <= 1e-9
    assert abs(f[0]["forces"]["total"][0, 2] - 100.0) <= 1e-9


def test_two_particle(oracle):
    sc, f = frames_of(oracle, "two_particle")
    c = f[0]["forces"]["contact"]
    assert abs(c[0, 0] + 100.0) <= 1e-9
    assert abs(c[1, 0] - 100.0) <= 1e-9
    assert abs(c[0, 0] + c[1, 0]) <= 1e-12


def regime_invariants(counts):
    """validateBodyRegimes (scenario_validator.cpp:32-73)."""
    for row in counts:
        active, stick, slide_p, twist_p, slide, twist, both = (int(x) for x in row)
        assert stick + slide_p + twist_p == active
        assert both <= slide and both <= twist
        assert slide + twist - both <= active
        assert slide_p == slide - both
        assert twist_p == twist
        assert stick == active - slide - twist + both


def test_triangle_contact(oracle):
    sc, f = frames_of(oracle, "triangle_contact")
    fp = f[0]["forces"]["total"][0, 2]
    fb = f[0]["extra"]["loads"][0, 2]
    assert abs(fp - 100.0) <= 1e-8
    assert abs(fb + 100.0) <= 1e-8
    assert abs(fp + fb) <= 1e-10
    saw = False
    for fr in f.values():
        regime_invariants(fr["stats"]["regime_instant"])
        regime_invariants(fr["stats"]["regime_interval"])
        saw = saw or fr["stats"]["regime_instant"][0, 0] > 0
    assert saw


def test_body_states(oracle):
    sc, f = frames_of(oracle, "body_states", interval=10)
    b = f[10]["extra"]["bodies"]
    assert abs(b["center"][0, 0]) <= 1e-12 and abs(b["vel"][0, 0]) <= 1e-12
    assert abs(b["center"][1, 0] - 10.02) <= 1e-11
    assert abs(b["vel"][1, 0] - 2.0) <= 1e-12
    assert abs(b["angvel"][1, 2] - 1.0) <= 1e-12
    assert abs(b["orien"][1, 3] - math.sin(0.005)) <= 2e-9


# ---- sphere / triangle area known answers (contact_geometry_test.cpp:37-76) --
def _area_fn():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    f = lib.ora_circle_triangle_area
    arr = C.c_double * 3
    f.argtypes = [arr, arr, arr, arr, arr, C.c_double]
    f.restype = C.c_double

    def sphere_area(center, radius, a, b, c):
        # sphere section: plane through a, b, c; circle radius from the distance
        a, b, c, center = (np.asarray(x, float) for x in (a, b, c, center))
        n = np.cross(b - a, c - a)
        ln = np.linalg.norm(n)
        if ln == 0.0:
            n = np.array([0.0, 0.0, 1.0])
        else:
            n = n / ln
        d = float(np.dot(center - a, n))
        r2 = radius * radius - d * d
        if r2 <= 0.0:
            return 0.0
        cc = center - n * d
        return f(arr(*cc), arr(*n), arr(*a), arr(*b), arr(*c), math.sqrt(r2))

    return sphere_area


@pytest.mark.parametrize("name,expected,center,radius,a,b,c", [
    ("disk inside triangle", math.pi, (0, 0, 0), 1.0, (-10, -10, 0), (10, -10, 0), (0, 10, 0)),
    ("offset disk inside triangle", 0.64 * math.pi, (0, 0, 0.6), 1.0, (-10, -10, 0), (10, -10, 0), (0, 10, 0)),
    ("sphere tangent to plane", 0.0, (0, 0, 1.0), 1.0, (-10, -10, 0), (10, -10, 0), (0, 10, 0)),
    ("sphere misses plane", 0.0, (0, 0, 2.0), 1.0, (-10, -10, 0), (10, -10, 0), (0, 10, 0)),
    ("triangle inside disk", 0.5, (0, 0, 0), 10.0, (0, 0, 0), (1, 0, 0), (0, 1, 0)),
    ("disjoint", 0.0, (0, 0, 0), 1.0, (2, 2, 0), (3, 2, 0), (2, 3, 0)),
    ("half disk", 0.5 * math.pi, (0, 0, 0), 1.0, (0, -10, 0), (10, 0, 0), (0, 10, 0)),
    ("quarter disk", 0.25 * math.pi, (0, 0, 0), 1.0, (0, 0, 0), (10, 0, 0), (0, 10, 0)),
    ("reversed orientation", 0.25 * math.pi, (0, 0, 0), 1.0, (0, 10, 0), (10, 0, 0), (0, 0, 0)),
    ("zero-radius sphere", 0.0, (0, 0, 0), 0.0, (-10, -10, 0), (10, -10, 0), (0, 10, 0)),
    ("large-scale invariance", math.pi * 1e12, (3e6, -2e6, 5e6), 1e6, (-7e6, -12e6, 5e6),
     (13e6, -12e6, 5e6), (3e6, 8e6, 5e6)),
])
def test_area_known_answers(name, expected, center, radius, a, b, c):
    area = _area_fn()(center, radius, a, b, c)
    scale = max(1.0, abs(area), abs(expected))
    assert abs(area - expected) <= 1e-10 * scale, (name, area, expected)


# ---- component transform (component_transform_test.cu:66-102) -----------
def test_component_offset_transform(oracle):
    """Owner at (3,4,5), q = 90 deg about z, body offset (1,0,0): the component
    sits at (3,5,5) and moves with v + (R w) x r = (0.1, 0.2, -0.7).  A wall at
    y = 5.05 (normal -y) touches it with depth 0.05; at the first evaluation
    (history dt = 0) the wall force is kN*depth + cN*v_n along -y and the
    friction is the dashpot on the contact-point slip: -(0.1, ., -0.7) plus
    the rolling term (n x w_world) * (r - depth) = (0, 0, 0.05)."""
    h = math.sqrt(0.5)
    P = dict(pid=np.array([0], np.int32), status=np.array([0], np.int32), mass=np.array([2.0]),
             radius=np.array([0.1]), inertia=np.array([[1.0, 1.0, 1.0]]), pos=np.array([[3.0, 4.0, 5.0]]),
             vel=np.array([[0.1, 0.2, 0.3]]), angvel=np.array([[0.0, 1.0, 0.0]]),
             quat=np.array([[h, 0.0, 0.0, h]]))
    eps, kN = 0.8, 1000.0
    m = scene.mech(epsS=eps, mu=10.0, epsN=eps, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=kN)
    sc = scene.Scene(**P, comp_id=np.array([0], np.int32), comp_owner=np.array([0], np.int32),
                     comp_radius=np.array([0.1]), comp_pos_body=np.array([[1.0, 0.0, 0.0]]),
                     wall_origin=np.array([[0.0, 5.05, 0.0]]), wall_normal=np.array([[0.0, -1.0, 0.0]]),
                     pp=m, pt=dict(m), pw=dict(m), dt=1e-4)
    s = oracle.solver()
    s.initialize(sc)
    s.compute_forces(collect=True)
    w = s.forces()["wall"][0]
    ln = math.log(eps)
    cN = -2 * ln * math.sqrt(kN * 2.0 / (math.pi ** 2 + ln * ln))
    cS = -2 * ln * math.sqrt((2 * kN / 7) * 2.0 / (math.pi ** 2 + ln * ln))
    depth = 0.05
    expected = np.array([-cS * 0.1, -(kN * depth + cN * 0.2), cS * (0.7 + 0.05)])
    assert abs(w[1] - expected[1]) <= 1e-12 * abs(expected[1]), (w, expected)
    assert abs(w[0] - expected[0]) <= 1e-12 and abs(w[2] - expected[2]) <= 1e-12, (w, expected)
    # torque = couple + offset x F with world offset (0,1,0)
    assert np.all(np.isfinite(w[3:]))


# ---- reference CPU binary at zero velocity (legacy law == CUDA law there) --
@pytest.mark.parametrize("tag", ["pure", "mbd"])
def test_data1_zero_velocity_matches_reference_cpu(oracle, tag):
    """DATA1 with every velocity zeroed: the legacy CPU law and the CUDA law
    coincide at the first force evaluation (no dashpot, no history), so the
    oracle's particle forces, body-to-particle forces and body loads must
    equal what the reference's CPU binary writes (to its 12 printed digits)."""
    z = np.load(os.path.join(GOLDEN, "data1_zero_vel_ref_cpu.npz"))
    sc = golden(f"data1_{tag}.npz")
    sc.vel = np.zeros_like(sc.vel)
    sc.angvel = np.zeros_like(sc.angvel)
    sc.keep_body_particle = 1
    s = oracle.solver()
    s.initialize(sc)
    loads = s.compute_forces(sc.body_pose if len(sc.body_mass) else None, loads=len(sc.body_mass) > 0,
                             collect=True)
    f = s.forces()
    F = z[f"{tag}__F_0"]
    mine = np.concatenate([f["contact"], f["wall"], f["field"][:, :3], f["total"]], axis=1)
    assert_close(mine, F[:, :21], 2e-11, 1e-10, f"{tag} initial forces vs reference CPU")
    if tag == "mbd":
        nb = len(sc.body_mass)
        fi = f["body_particle"].transpose(1, 0, 2).reshape(len(F), 6 * nb)
        assert_close(fi, F[:, 21:], 2e-11, 1e-10, "FI_pt vs reference CPU")
        At = z["mbd__At_0"]
        assert_close(loads, At[:, 14:20], 2e-11, 1e-9, "body loads vs reference CPU")


# ---- mechanics ------------------------------------------------------------
def frozen_pair(mu, vt, dt=1e-5, kN=1000.0, overlap=1e-4, eps=0.9):
    """Two frozen spheres (status -1: positions fixed) with a constant
    tangential relative velocity: the tangential spring accumulates
    kS * vt * dt per step up to the Coulomb limit mu * |Fn|."""
    r = 0.01
    P = scene.sphere_particles([[0.0, 0.0, 0.0], [2 * r - overlap, 0.0, 0.0]], [r, r], mass=[1.0, 1.0],
                               status=[-1, -1])
    P["vel"] = np.array([[0.0, 0.0, 0.0], [0.0, vt, 0.0]])
    m = scene.mech(epsS=eps, mu=mu, epsN=eps, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=kN)
    return scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=dt)


def test_spring_slider_accumulates_then_caps(oracle):
    kN, overlap, dt, vt, mu, eps = 1000.0, 1e-4, 1e-5, 1.0, 0.4, 0.9
    sc = frozen_pair(mu, vt, dt, kN, overlap, eps)
    s = oracle.solver()
    s.initialize(sc)
    kS = 2 * kN / 7
    ln = math.log(eps)
    cS = -2 * ln * math.sqrt(kS * 0.5 / (math.pi ** 2 + ln * ln))
    fn = kN * overlap
    forces = []
    s.compute_forces(collect=True)
    forces.append(s.forces()["contact"][0, 1])
    s.init_half_step(0.5 * dt)
    for _ in range(60):
        s.advance(dt)
        s.compute_forces(collect=True)
        s.finish(dt)
        forces.append(s.forces()["contact"][0, 1])
    # step 0: history dt = 0, only the dashpot acts
    assert abs(forces[0] - cS * vt) <= 1e-12
    for k in range(1, 61):
        spring = kS * vt * dt * k
        if spring <= mu * fn * (1 - 1e-9):
            assert abs(forces[k] - (spring + cS * vt)) <= 1e-12 * max(1.0, abs(forces[k])), k
    # well past the cap the friction is the Coulomb limit along the slip
    assert abs(forces[-1] - mu * fn) <= 1e-12
    assert kS * vt * dt * 60 > mu * fn


def test_normal_restitution(oracle):
    """Head-on collision, no friction, linear spring-dashpot: the outgoing to
    incoming normal relative speed ratio is the restitution coefficient."""
    eps, kN, r = 0.6, 1e4, 1e-3
    P = scene.sphere_particles([[0.0, 0.0, 0.0], [2 * r + 1e-6, 0.0, 0.0]], [r, r], density=2500.0)
    P["vel"] = np.array([[0.05, 0.0, 0.0], [-0.05, 0.0, 0.0]])
    m = scene.mech(epsS=eps, mu=0.0, epsN=eps, cohesion=0.0, beta=0.5, muT=0.0, muR=0.0, kN=kN)
    dt = 2e-8
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=dt, steps=12000)
    s = oracle.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=12000)
    v = s.state()["vel"]
    ratio = (v[1, 0] - v[0, 0]) / 0.1
    assert abs(ratio - eps) <= 2e-3, ratio


def test_action_reaction_in_a_box(oracle):
    """Pair forces balance (momentum) and pair torques balance about the
    origin (angular momentum) to rounding, after 30 steps of a jittered
    overlapping lattice with random linear and angular velocities."""
    rng = np.random.RandomState(1)
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3) * 1.9e-3
    pos = g + rng.uniform(-5e-5, 5e-5, g.shape)
    P = scene.sphere_particles(pos, rng.uniform(0.98e-3, 1.02e-3, len(g)), density=2500.0)
    P["vel"] = rng.uniform(-0.2, 0.2, (len(g), 3))
    P["angvel"] = rng.uniform(-40, 40, (len(g), 3))
    m = scene.mech(epsS=0.5, mu=0.4, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.2, muR=0.1, kN=1.25e4)
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=1e-6, steps=30)
    s = oracle.solver()
    s.initialize(sc)
    host.run_pure(s, sc, steps=30)
    s.compute_forces(collect=True)
    c = s.forces()["contact"]
    assert s.stats()["particle_contacts"] > 2000
    F = c[:, :3]
    scale = np.abs(F).sum()
    assert np.abs(F.sum(axis=0)).max() <= 1e-12 * scale
    x = s.state()["pos"]
    L = np.cross(x, F).sum(axis=0) + c[:, 3:].sum(axis=0)
    assert np.abs(L).max() <= 1e-11 * (np.abs(np.cross(x, F)).sum() + np.abs(c[:, 3:]).sum())


def test_cell_key_order_is_lexicographic():
    """packCell (grid_backend.cu:19-26): the 64-bit key orders cells by x, then
    y, then z over the signed 21-bit range."""
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.ora_pack_cell.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ora_pack_cell.restype = C.c_uint64
    rng = np.random.RandomState(0)
    lim = 1 << 20
    cells = [tuple(int(v) for v in rng.randint(-lim, lim, 3)) for _ in range(400)]
    cells += [(-lim, -lim, -lim), (lim - 1, lim - 1, lim - 1), (0, 0, 0), (-1, 0, 0), (0, -1, 0)]
    keys = [lib.ora_pack_cell(*c) for c in cells]
    assert [c for _, c in sorted(zip(keys, cells))] == sorted(cells)
    assert lib.ora_pack_cell(-lim, -lim, -lim) == 0
    assert lib.ora_pack_cell(0, 0, 0) == (lim << 42) | (lim << 21) | lim


def test_grid_range_error_message(oracle):
    sc = scene.box_scene(n=50, seed=2, box=(0.02, 0.02, 0.02))
    sc.pos[7, 0] = float("nan")
    s = oracle.solver()
    s.initialize(sc)
    with pytest.raises(Exception, match="component grid coordinate exceeds signed 21-bit hash range"):
        s.compute_forces(collect=True)


def test_history_capacity_error(oracle):
    """A large sphere touching more than 64 small ones overflows the particle
    history; the diagnostics step raises the reference's message."""
    rng = np.random.RandomState(0)
    big = 0.01
    dirs = rng.normal(size=(90, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    small = 0.0004
    pos = np.vstack([[0.0, 0.0, 0.0], dirs * (big + small * 0.5)])
    rad = np.concatenate([[big], np.full(90, small)])
    P = scene.sphere_particles(pos, rad, density=2500.0)
    m = scene.mech(epsS=0.5, mu=0.5, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=100.0)
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=1e-6)
    s = oracle.solver()
    s.initialize(sc)
    with pytest.raises(Exception, match="CUDA particle contact-history capacity exceeded"):
        s.compute_forces(collect=True)


def test_arc_tangent_accuracy():
    """The area routine's atan2 (half-angle reductions + Taylor series, the
    same evaluation on CPU and GPU) against the C library's atan2: within 3
    ulp over random and axis-adjacent arguments, exact special cases."""
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    f = lib.ora_atan2
    f.argtypes = [C.c_double, C.c_double]
    f.restype = C.c_double
    rng = np.random.RandomState(7)
    ys = np.concatenate([rng.normal(size=20000), rng.normal(size=2000) * 1e-9, [1.0, -1.0, 1e-300]])
    xs = np.concatenate([rng.normal(size=20000), rng.normal(size=2000), [1e-300, -1e-300, 1.0]])
    for y, x in zip(ys, xs):
        ref = math.atan2(y, x)
        got = f(y, x)
        assert abs(got - ref) <= 3 * math.ulp(ref), (y, x, got, ref)
    for y, x in [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0),
                 (0.0, 0.0), (0.0, -0.0)]:
        assert f(y, x) == math.atan2(y, x) and math.copysign(1, f(y, x)) == math.copysign(1, math.atan2(y, x))


def test_oracle_threads_do_not_change_results(oracle):
    """The oracle's owner loop runs on several threads; one thread and
    eight give the same bits (states, forces, history tables, statistics)
    on a box of 3,000 agitated spheres over 20 steps."""
    from dem_mbd_cuda_amd import host

    sc = scene.box_scene(n=3000, seed=23, box=(0.05, 0.05, 0.06), steps=20, dt=1e-6)
    rng = np.random.RandomState(24)
    sc.vel = rng.uniform(-0.5, 0.5, (sc.n, 3))
    sc.angvel = rng.uniform(-40, 40, (sc.n, 3))
    out = []
    for threads in (1, 8):
        s = oracle.solver()
        s.set_threads(threads)
        s.initialize(sc)
        host.run_pure(s, sc)
        out.append((s.state(), s.forces(), s.debug_history(), s.stats()))
        s.close()
    (a, fa, ha, sa), (b, fb, hb, sb) = out
    assert sa["particle_contacts"] > 0 and sa["particle_contacts"] == sb["particle_contacts"]
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for k in ("contact", "wall", "total"):
        assert np.array_equal(fa[k], fb[k]), k
    for k in ha:
        assert np.array_equal(ha[k], hb[k]), k
