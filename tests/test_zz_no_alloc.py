"""dem_compute_forces_noalloc, whose triangle grid lives in the device's static
scratch array, against dem_compute_forces: every output of a coupled run bit
for bit, for two scenes run one after the other on the same device (the
second with a smaller triangle grid, over what the first left in the array)
and the first again after it; a step whose triangle grid exceeds the array
fails with DEM_E_CAPACITY and a message naming the limit."""
import numpy as np
import pytest

from conftest import golden
from dem_mbd_cuda_amd import host, scene
from dem_mbd_cuda_amd._abi import DemError

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def coupled_frames(lib, sc, noalloc, steps):
    s = lib.solver()
    s.noalloc = noalloc
    s.initialize(sc)
    frames = {}

    def cb(step, t, extra):
        frames[step + 1] = dict(state=s.state(), forces=s.forces(), stats=s.stats(), extra=extra,
                                history=s.debug_history(), tgrid=s.debug_triangle_grid(),
                                entries=s.grid_entry_counts()[1])

    host.run_coupled(s, sc, steps=steps, output_interval=1, on_frame=cb)
    s.close()
    return frames


def assert_same(a, b):
    assert sorted(a) == sorted(b)
    for i in sorted(a):
        fa, fb = a[i], b[i]
        for k in ("pos", "vel", "angvel", "quat"):
            assert np.array_equal(fa["state"][k], fb["state"][k]), (i, k)
        for k in ("contact", "field", "wall", "total", "body_particle", "body_detail"):
            if fb["forces"][k] is not None:
                assert np.array_equal(fa["forces"][k], fb["forces"][k]), (i, k)
        for k in ("particle_contacts", "body_contacts", "body_max_depth", "body_contact_fz"):
            assert fa["stats"][k] == fb["stats"][k], (i, k)
        for k in ("regime_instant", "regime_interval", "micro_instant", "micro_interval"):
            assert np.array_equal(fa["stats"][k], fb["stats"][k]), (i, k)
        for k in fb["history"]:
            assert np.array_equal(fa["history"][k], fb["history"][k]), (i, k)
        assert np.array_equal(fa["tgrid"][0], fb["tgrid"][0]) and np.array_equal(fa["tgrid"][1], fb["tgrid"][1]), i
        assert np.array_equal(fa["extra"]["loads"], fb["extra"]["loads"]), i
        for k in ("center", "orien", "vel", "angvel"):
            assert np.array_equal(fa["extra"]["bodies"][k], fb["extra"]["bodies"][k]), (i, k)


def test_noalloc_matches_compute_forces_on_two_scenes_in_a_row(hip):
    big = scene.mixer_scene(n=20_000, packing="dense", steps=6)
    big.keep_body_particle = 1
    big.keep_body_detail = 1
    small = golden("case_triangle_contact.npz")
    small.keep_body_particle = 1
    small.keep_body_detail = 1
    ref_big = coupled_frames(hip, big, False, 6)
    ref_small = coupled_frames(hip, small, False, None)
    print("triangle grid entries: mixer", ref_big[0]["entries"], "small case", ref_small[0]["entries"], flush=True)
    assert ref_big[0]["entries"] > 10 * max(1, ref_small[0]["entries"])
    assert ref_big[1]["stats"]["body_contacts"] > 0
    na_big = coupled_frames(hip, big, True, 6)
    assert_same(na_big, ref_big)
    # second call, smaller grid, over the mixer's contents of the array
    na_small = coupled_frames(hip, small, True, None)
    assert_same(na_small, ref_small)
    # and the larger scene again after it
    assert_same(coupled_frames(hip, big, True, 6), ref_big)


def test_noalloc_rejects_a_grid_beyond_the_scratch_array(hip):
    """A 25 mm triangle over 0.05 mm spheres: its bounding box (with the
    inner apex 28 mm below it) spans about 201 x 201 x 229 cells of 0.125 mm,
    9.3M entries in a dense table of as many cells, about 150 MB of grid
    buffers, beyond the 32 MiB array."""
    P = scene.sphere_particles([[0.0, 0.0, 1e-3], [0.005, 0.0, 1e-3]], [5e-5, 5e-5], density=2500.0)
    rows = scene.rectangle_mesh((-0.0125, -0.0125, 0.0), (0.025, 0, 0), (0, 0.025, 0), 1, 1)[:1]
    m = scene.mech(epsS=0.5, mu=0.4, epsN=0.5, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=2.0e4)
    sc = scene.Scene(**P, pp=m, pt=dict(m), pw=dict(m), dt=1e-6, steps=0, body_mass=np.array([1.0]),
                     body_pose=np.array([[0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0]], float),
                     body_tri_count=np.array([1], np.int32), tri_local=rows,
                     body_state=np.array([1], np.int32), body_inertia=np.eye(3)[None])
    s = hip.solver()
    s.noalloc = True
    s.initialize(sc)
    with pytest.raises(DemError) as e:
        s.compute_forces(body_pose=sc.body_pose, loads=True, collect=True)
    msg = str(e.value)
    assert "no-allocation scratch too small" in msg and str(SCRATCH_BYTES) in msg, msg
    s.close()
