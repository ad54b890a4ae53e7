"""HIP path (libdemhip.so through the C-ABI) against the CPU oracle on the
same inputs.

Both sides evaluate the law without FMA contraction, with correctly rounded
+, -, *, /, sqrt, in the same expression order, the same arc-tangent
evaluation and the same fixed reduction tree for body loads, so every value
is expected to agree bit for bit: cell keys, sorted orders and history
tables (the contact-pair sets) exactly, and states, forces, history values,
body loads, regime counters and microstate sums to the last bit.
"""
import numpy as np
import pytest

from conftest import golden
from helpers import assert_close, run_frames

pytestmark = pytest.mark.gpu

CASES = ["free_fall", "wall_contact", "two_particle", "triangle_contact", "body_states"]


def compare_frames(g, o, tol=0.0):
    """tol = 0: bit-exact (numerical equality); otherwise relative tolerance."""
    atol = 0.0 if tol == 0.0 else 1e-300
    assert sorted(g) == sorted(o)
    for idx in sorted(o):
        gs, os_ = g[idx]["state"], o[idx]["state"]
        for k in ("pos", "vel", "angvel", "quat"):
            assert_close(gs[k], os_[k], tol, atol, f"frame {idx} {k}")
        for k in ("contact", "wall", "field", "total"):
            assert_close(g[idx]["forces"][k], o[idx]["forces"][k], tol, atol, f"frame {idx} force {k}")
        for k in ("body_particle", "body_detail"):
            if o[idx]["forces"][k] is not None:
                assert_close(g[idx]["forces"][k], o[idx]["forces"][k], tol, atol, f"frame {idx} {k}")
        gst, ost = g[idx]["stats"], o[idx]["stats"]
        for k in ("particle_contacts", "body_contacts", "particle_history_overflows",
                  "body_history_overflows", "particle_history_high_water", "body_history_high_water"):
            assert gst[k] == ost[k], (idx, k, gst[k], ost[k])
        for k in ("body_max_depth", "body_contact_fz"):
            assert_close(gst[k], ost[k], tol, atol, f"frame {idx} {k}")
        for k in ("regime_instant", "regime_interval"):
            assert np.array_equal(gst[k], ost[k]), (idx, k)
        for k in ("micro_instant", "micro_interval"):
            assert_close(gst[k], ost[k], tol, atol, f"frame {idx} {k}")
        if "grid" in o[idx]:
            gk, gi = g[idx]["grid"]
            ok, oi = o[idx]["grid"]
            assert np.array_equal(gk, ok), f"frame {idx}: cell keys differ"
            assert np.array_equal(gi, oi), f"frame {idx}: sorted order differs"
            gh, oh = g[idx]["history"], o[idx]["history"]
            for k in ("pp_keys", "pw_keys", "pt_keys"):
                assert np.array_equal(gh[k], oh[k]), f"frame {idx}: {k} (contact-pair set) differs"
            for k in ("pp_values", "pw_values", "pt_values"):
                assert_close(gh[k], oh[k], tol, atol, f"frame {idx} {k}")
        if o[idx]["extra"] is not None:
            ge, oe = g[idx]["extra"], o[idx]["extra"]
            assert_close(ge["loads"], oe["loads"], tol, atol, f"frame {idx} body loads")
            for k in ("center", "orien", "vel", "angvel"):
                assert_close(ge["bodies"][k], oe["bodies"][k], tol, atol, f"frame {idx} body {k}")
        if "tgrid" in o[idx]:
            gk, gt = g[idx]["tgrid"]
            ok, ot = o[idx]["tgrid"]
            assert np.array_equal(gk, ok) and np.array_equal(gt, ot), "triangle grid differs"


@pytest.mark.parametrize("case", CASES)
def test_validation_case(hip, oracle, case):
    sc = golden(f"case_{case}.npz")
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    g = run_frames(hip, sc, interval=1, debug=True)
    o = run_frames(oracle, sc, interval=1, debug=True)
    compare_frames(g, o)


def test_box_bit_exact(hip, oracle):
    """2,000-sphere box (radius U[1,1.5] mm, 6 walls, gravity, seeded RSA),
    150 steps with collisions enabled by an initial velocity field."""
    from dem_mbd_cuda_amd import scene

    sc = scene.box_scene(n=2000, seed=5, steps=150, dt=1e-6)
    rng = np.random.RandomState(11)
    sc.vel = rng.uniform(-0.5, 0.5, (sc.n, 3))
    sc.angvel = rng.uniform(-50, 50, (sc.n, 3))
    g = run_frames(hip, sc, interval=50, debug=True)
    o = run_frames(oracle, sc, interval=50, debug=True)
    assert o[150]["stats"]["particle_contacts"] > 0
    compare_frames(g, o)


@pytest.mark.parametrize("reorder", ["0", "1", "7"])
def test_reorder_interval_invariance(hip, oracle, reorder, monkeypatch):
    """Physically reordering the particle arrays must not change any result."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_REORDER_INTERVAL", reorder)
    sc = scene.box_scene(n=800, seed=9, box=(0.04, 0.04, 0.04), steps=40)
    rng = np.random.RandomState(3)
    sc.vel = rng.uniform(-0.3, 0.3, (sc.n, 3))
    g = run_frames(hip, sc, interval=20, debug=True)
    o = run_frames(oracle, sc, interval=20, debug=True)
    compare_frames(g, o)


def test_candidate_lists_rebuilt_every_few_steps(hip, oracle, monkeypatch):
    """A skin of 0.001 x the largest radius (the rebuild threshold, half the
    skin, is 0.75 um) with spheres moving about 0.5 um per step: candidate
    lists are rebuilt every few steps, so list steps whose queued scan runs
    alternate with steps whose queued scan finds the rebuild flag and does
    nothing, and the rebuild path takes over."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_SKIN", "0.001")
    sc = scene.box_scene(n=1500, seed=13, box=(0.04, 0.04, 0.04), steps=60, dt=1e-6)
    rng = np.random.RandomState(17)
    sc.vel = rng.uniform(-0.5, 0.5, (sc.n, 3))
    sc.angvel = rng.uniform(-30, 30, (sc.n, 3))
    g = run_frames(hip, sc, interval=20, debug=True)
    o = run_frames(oracle, sc, interval=20, debug=True)
    assert o[60]["stats"]["particle_contacts"] > 0
    compare_frames(g, o)


def test_data1_coupled(hip, oracle):
    """DATA1 (1,704 spheres, 5 walls, 2 dynamic bodies x 1,416 triangles, host
    Newton-Euler between steps), 5 steps, bit for bit."""
    sc = golden("data1_mbd.npz")
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    g = run_frames(hip, sc, interval=1, debug=True)
    o = run_frames(oracle, sc, interval=1, debug=True)
    compare_frames(g, o)


def test_single_contact_1000_steps(hip, oracle):
    """A sphere resting on a frozen sphere (status -1) under gravity, sliding,
    spinning and rolling on it for 1,000 steps: one persistent contact pair.
    Positions, velocities and forces within 1e-10 relative of the oracle."""
    from dem_mbd_cuda_amd import scene

    r, rho, kN, g = 1e-3, 2500.0, 1.25e4, 9.81
    m = rho * 4.0 / 3.0 * np.pi * r ** 3
    overlap = m * g / kN
    P = scene.sphere_particles([[0.0, 0.0, 0.0], [0.0, 0.0, 2 * r - overlap]], [r, r], density=rho,
                               status=[-1, 0])
    P["vel"] = np.array([[0.0, 0.0, 0.0], [0.02, 0.005, 0.0]])
    P["angvel"] = np.array([[0.0, 0.0, 0.0], [5.0, 30.0, 20.0]])
    mm = scene.mech(epsS=0.5, mu=0.3, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=kN)
    sc = scene.Scene(**P, pp=mm, pt=dict(mm), pw=dict(mm), dt=1e-6, steps=1000, gravity=(0.0, 0.0, -g),
                     gravity_enabled=1)
    g_ = run_frames(hip, sc, interval=100)
    o = run_frames(oracle, sc, interval=100)
    for idx in o:
        assert o[idx]["stats"]["particle_contacts"] == 2, idx
    compare_frames(g_, o, tol=1e-10)


@pytest.mark.parametrize("skin", ["0.2", "0"])
def test_mixer_with_all_three_body_states(hip, oracle, skin, monkeypatch):
    """A small mixer (3,000 spheres in a triangle-mesh bin, fixed; a blade in
    prescribed rotation; a dynamic plate pressed onto the bed), seeded
    velocities and spins, 60 coupled steps: particles, body kinematics,
    loads, regime counters and microstate sums bit for bit; with candidate
    lists (their rebuilds move the triangle slots with the rows) and with
    the full scan every step (DEM_SKIN=0)."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_SKIN", skin)

    sc = scene.mixer_scene(n=3000, side=0.05, seed=8, dt=1e-5, blade_clear=0.0, plate_gap=-0.3e-3)
    rng = np.random.RandomState(9)
    sc.vel = rng.uniform(-0.3, 0.3, (sc.n, 3))
    sc.vel[:, 2] -= 0.5
    sc.angvel = rng.uniform(-50, 50, (sc.n, 3))
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    g = run_frames(hip, sc, steps=60, interval=20, debug=True)
    o = run_frames(oracle, sc, steps=60, interval=20, debug=True)
    assert o[0]["stats"]["body_contacts"] > 0 and o[60]["stats"]["particle_contacts"] > 0
    compare_frames(g, o)


def test_dense_mixer_bed_pressed_onto_every_body(hip, oracle):
    """The settled-bed mixer (4,000 spheres on the bin floor and walls, on
    both faces of the turning blade and under the plate), 60 coupled steps:
    the layout the coupled throughput figure runs, bit for bit."""
    from dem_mbd_cuda_amd import scene

    sc = scene.mixer_scene(n=4000, side=0.05, seed=10, dt=1e-5, packing="dense")
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    g = run_frames(hip, sc, steps=60, interval=20, debug=True)
    o = run_frames(oracle, sc, steps=60, interval=20, debug=True)
    assert o[0]["stats"]["body_contacts"] > 0.1 * sc.n
    compare_frames(g, o)



def test_triangle_grid_box_beyond_the_cell_table(hip, oracle):
    """A fixed plate and a dynamic lid 2 m apart: the triangle grid's cell
    box (about 800^3 cells of 2.5 mm) is too large for the dense cell table,
    so the grid is built as sorted (key, triangle) pairs and searched.
    Spheres start pressed into the plate and the lid comes down onto a
    sphere below it; 40 coupled steps, bit for bit, including the grid."""
    from dem_mbd_cuda_amd import scene

    rng = np.random.RandomState(13)
    plate = scene.rectangle_mesh((-0.01, -0.01, 0.0), (0.02, 0, 0), (0, 0.02, 0), 4, 4)  # faces up
    lid = scene.rectangle_mesh((-0.005, -0.005, 0.0), (0, 0.01, 0), (0.01, 0, 0), 2, 2)  # faces down
    # 15 spheres pressed 0.05 mm into the first plate, one 0.05 mm below the lid
    pos = np.column_stack([rng.uniform(-0.008, 0.008, 16), rng.uniform(-0.008, 0.008, 16), np.full(16, 0.95e-3)])
    pos[15] = [2.0, 2.0, 2.0 - 1.05e-3]
    P = scene.sphere_particles(pos, np.full(16, 1e-3), density=2500.0)
    m = scene.mech(epsS=0.5, mu=0.4, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=2.0e4)
    ident = [1.0, 0.0, 0.0, 0.0]
    pose = np.array([[0.0, 0.0, 0.0, *ident, 0, 0, 0, 0, 0, 0],
                     [2.0, 2.0, 2.0, *ident, 0, 0, -0.2, 0, 0, 0]])
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=1e-5,
                     gravity=(0.0, 0.0, -9.81), gravity_enabled=1, steps=40,
                     body_mass=np.array([100.0, 0.01]), body_pose=pose,
                     body_tri_count=np.array([len(plate), len(lid)], np.int32),
                     tri_local=np.concatenate([plate, lid]), body_state=np.array([1, 0], np.int32),
                     body_inertia=np.stack([np.eye(3), np.eye(3) * 1e-6]))
    g = run_frames(hip, sc, steps=40, interval=10, debug=True)
    o = run_frames(oracle, sc, steps=40, interval=10, debug=True)
    assert o[0]["stats"]["body_contacts"] > 0
    compare_frames(g, o)


def test_universal_gravitation_single_spheres(hip, oracle):
    """Universal gravitation on single-sphere owners (the O(N^2) field sum in
    input-row order, compiled into the law kernel only for such runs) with
    walls and contacts, 40 steps, bit for bit."""
    from dem_mbd_cuda_amd import scene

    sc = scene.box_scene(n=400, seed=17, box=(0.025, 0.025, 0.04), steps=40, dt=1e-6)
    sc.universal_gravity = 1
    sc.vel = np.random.RandomState(18).uniform(-2.0, 2.0, (sc.n, 3))
    g = run_frames(hip, sc, interval=20, debug=True)
    o = run_frames(oracle, sc, interval=20, debug=True)
    assert o[40]["stats"]["particle_contacts"] > 0
    compare_frames(g, o)


def test_fine_mesh_rows_beyond_the_stored_hits(hip, oracle):
    """Spheres pressed into a fine triangle floor (0.25 mm triangles): every
    cell holds hundreds of triangle entries and a sphere touches about 20
    triangles, more than the per-row hit store keeps, so the contacts take
    the in-lane path; 30 coupled steps, bit for bit."""
    from dem_mbd_cuda_amd import scene

    rng = np.random.RandomState(21)
    floor = scene.rectangle_mesh((-0.01, -0.01, 0.0), (0.02, 0, 0), (0, 0.02, 0), 80, 80)  # faces up
    pos = np.column_stack([rng.uniform(-0.008, 0.008, 40), rng.uniform(-0.008, 0.008, 40),
                           rng.uniform(0.85e-3, 0.95e-3, 40)])
    P = scene.sphere_particles(pos, np.full(40, 1e-3), density=2500.0)
    P["vel"] = rng.uniform(-0.2, 0.2, (40, 3))
    m = scene.mech(epsS=0.5, mu=0.4, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=2.0e4)
    pose = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0]])
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=1e-5, gravity=(0.0, 0.0, -9.81), gravity_enabled=1,
                     steps=30, body_mass=np.array([1.0]), body_pose=pose,
                     body_tri_count=np.array([len(floor)], np.int32), tri_local=floor,
                     body_state=np.array([1], np.int32), body_inertia=np.eye(3)[None])
    sc.keep_body_particle = 1
    sc.keep_body_detail = 1
    g = run_frames(hip, sc, steps=30, interval=10, debug=True)
    o = run_frames(oracle, sc, steps=30, interval=10, debug=True)
    assert o[0]["stats"]["body_contacts"] > 16 * 20
    compare_frames(g, o)


@pytest.mark.parametrize("skin", ["0", "0.01"])
def test_device_sort_order(hip, oracle, skin, monkeypatch):
    """The grid's sorted order as the device sorts it: 3,000 spheres in a
    box with several spheres per cell (ties ordered by input row), agitated,
    a frame every step.  With the full scan (DEM_SKIN=0) every evaluation
    sorts on the device; with a small skin the candidate lists are rebuilt
    every few steps.  On every frame whose evaluation sorted, dem_debug_grid
    returns the device's keys and order untouched, and they equal the
    oracle's (key, index) order bit for bit."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_SKIN", skin)
    sc = scene.box_scene(n=3000, seed=29, box=(0.05, 0.05, 0.06), steps=30, dt=1e-6, rmin=0.4e-3, rmax=1.5e-3)
    rng = np.random.RandomState(30)
    sc.vel = rng.uniform(-1.0, 1.0, (sc.n, 3))
    g = run_frames(hip, sc, interval=1, debug=True)
    o = run_frames(oracle, sc, interval=1, debug=True)
    sorted_frames = [i for i in g if g[i]["counters"]["last_sorted"]]
    if skin == "0":
        assert len(sorted_frames) == len(g)
    else:
        assert 3 <= len(sorted_frames) < len(g), sorted_frames
    keys = o[0]["grid"][0]
    assert len(np.unique(keys)) < 0.9 * len(keys)  # ties exist
    compare_frames(g, o)


def _bimodal_scene():
    """Small spheres (r 0.5 mm) on a slightly overlapping cubic lattice under
    two large ones (r 4 mm): cells of 2 x 4 mm hold about 500 small spheres,
    so a block's staged neighbourhood exceeds the LDS stage."""
    from dem_mbd_cuda_amd import scene

    r, a = 0.5e-3, 0.999e-3
    g = np.arange(20) * a + r
    xs, ys, zs = np.meshgrid(g, g, g[:10], indexing="ij")
    small = np.stack([xs.ravel(), ys.ravel(), zs.ravel()], axis=1)
    big = np.array([[0.006, 0.006, 0.030], [0.015, 0.015, 0.030]])
    pos = np.vstack([small, big])
    radius = np.concatenate([np.full(len(small), r), np.full(2, 4e-3)])
    perm = np.random.RandomState(51).permutation(len(pos))
    pos, radius = pos[perm].copy(), radius[perm].copy()
    m = scene.mech(epsS=0.5, mu=0.5, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=1e4)
    wo, wn = scene.box_walls(np.zeros(3), np.array([0.021, 0.021, 0.04]))
    sc = scene.Scene(**scene.sphere_particles(pos, radius, 2500.0), wall_origin=wo, wall_normal=wn, pp=m,
                     pt=dict(m), pw=dict(m), dt=1e-6, gravity=(0.0, 0.0, -9.81), gravity_enabled=1,
                     mesh_scale=2.0, steps=10)
    sc.vel = np.random.RandomState(52).uniform(-0.05, 0.05, (sc.n, 3))
    return sc


@pytest.mark.parametrize("case", ["dense_bed", "bimodal", "dilute_box"])
def test_full_scan_staging_paths(hip, oracle, case, monkeypatch):
    """The full 27-cell scan (DEM_SKIN=0) on its three paths, bit for bit
    against the oracle: a settled hopper bed at a cell of 2 x max radius
    (the blocks' neighbourhoods staged in LDS), a bimodal packing whose
    staged neighbourhoods exceed the stage, and a dilute box whose blocks
    span more columns than are staged (both read the candidates from device
    memory).  The run counters say which path the diagnostics evaluations
    took."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.setenv("DEM_SKIN", "0")
    if case == "dense_bed":
        sc = scene.hopper_scene(n=20_000, width=0.2 * 0.02 ** 0.5, packing="settled", mesh_scale=2.0, steps=10)
    elif case == "bimodal":
        sc = _bimodal_scene()
    else:
        sc = scene.box_scene(n=2000, seed=53, steps=10)
        sc.vel = np.random.RandomState(54).uniform(-1.0, 1.0, (sc.n, 3))
    g = run_frames(hip, sc, steps=10, interval=5, debug=True)
    o = run_frames(oracle, sc, steps=10, interval=5, debug=True)
    c = g[10]["counters"]
    print(case, "scan passes staged / from device memory:", c["scan_lds_passes"], c["scan_hbm_passes"], flush=True)
    assert o[10]["stats"]["particle_contacts"] > 0
    if case == "dense_bed":
        # a bed this narrow has short columns, so some blocks span more
        # columns than are staged (about a quarter; 3 % on the 1M bed)
        assert c["scan_lds_passes"] > 2 * c["scan_hbm_passes"], c
    else:
        assert c["scan_hbm_passes"] > c["scan_lds_passes"], c
    compare_frames(g, o)


def test_adaptive_skin(hip, oracle, monkeypatch):
    """The candidate-list skin follows the motion (no DEM_SKIN given): spheres
    falling from rest under a strong field move slowly at first, so the
    first check shrinks the skin and rebuilds; as they speed up, rebuilds come
    sooner and the skin widens again.  Every frame equals the oracle's bit
    for bit, whatever the skin (any skin lists a superset of the contacts)."""
    from dem_mbd_cuda_amd import scene

    monkeypatch.delenv("DEM_SKIN", raising=False)
    monkeypatch.delenv("DEM_SKIN_ADAPT", raising=False)
    # g = 500 m/s^2: 1 um of fall in the first 64 steps (the first check),
    # then rebuilds every ~200 steps as the spheres speed up
    sc = scene.box_scene(n=2000, seed=61, steps=500, gravity=(0.0, 0.0, -500.0))
    sc.vel = np.random.RandomState(62).uniform(-0.005, 0.005, (sc.n, 3))
    g = run_frames(hip, sc, steps=500, interval=25, debug=False)
    o = run_frames(oracle, sc, steps=500, interval=25, debug=False)
    skins = [g[k]["counters"]["skin_pm"] for k in sorted(g)]
    rebuilds = g[max(g)]["counters"]["rebuilds"]
    print("skin (pm) per frame:", skins, "rebuilds", rebuilds, flush=True)
    start = skins[0]
    assert min(skins) <= start // 4, skins      # shrunk while the spheres barely moved
    assert skins[-1] >= 2 * min(skins), skins   # widened as they sped up
    assert rebuilds >= 3  # the first evaluation, the shrink, at least one widening
    compare_frames(g, o)
