"""The C-ABI library loads and exports every entry point include/dem_hip.h
declares; error behaviour that needs no device (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dem_hip.h")
LIB = os.path.join(ROOT, "dem_mbd_cuda_amd", "libdemhip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dem_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("dem_create", "dem_destroy", "dem_initialize", "dem_compute_forces",
                     "dem_init_half_step", "dem_advance_particles", "dem_finish_particle_step",
                     "dem_load_contact_history", "dem_save_contact_history",
                     "dem_has_restored_leapfrog_state", "dem_download_particle_state",
                     "dem_download_force_output", "dem_get_step_stats",
                     "dem_particle_grid_entry_count", "dem_triangle_grid_entry_count",
                     "dem_device_memory_bytes", "dem_backend_available", "dem_device_description",
                     "dem_last_error"):
        assert required in names, required


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libdemhip.so is not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_uninitialised_solver_errors():
    lib = C.CDLL(LIB)
    lib.dem_create.argtypes = [C.POINTER(C.c_void_p)]
    lib.dem_compute_forces.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    lib.dem_last_error.restype = C.c_char_p
    lib.dem_destroy.argtypes = [C.c_void_p]
    h = C.c_void_p()
    assert lib.dem_create(C.byref(h)) == 0
    rc = lib.dem_compute_forces(h, None, None, 0)
    assert rc != 0
    assert lib.dem_last_error().decode() == "CUDA DEM solver is not initialized"
    lib.dem_destroy(h)


def test_reference_host_links_against_the_adapter():
    """oracle/_ref/dem_mbd_hip is the reference's own host (CUDA-mode sources)
    linked to the adapter + libdemhip.so; --cuda-info goes through
    CudaBackendAvailable/CudaDeviceDescription of the adapter."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_hip")
    if not os.path.exists(exe):
        pytest.skip("reference host build not present (built only where the reference tree exists)")
    res = subprocess.run([exe, "--cuda-info"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    lib = C.CDLL(LIB)
    if lib.dem_backend_available():
        assert res.returncode == 0 and "CUDA device" in res.stdout, res.stdout
    else:
        assert res.returncode == 1 and "No usable NVIDIA CUDA device" in res.stdout, res.stdout


def _small_batch():
    from dem_mbd_cuda_amd import scene

    scenes = scene.sweep_scenes(count=3, seed=3, nmin=50, nmax=120)
    return scene, scenes


def test_batch_argument_errors_need_no_device():
    """dem_initialize_batch validates the scene layout before touching a
    device: scene ranges must cover the views, scenes must be non-empty, and
    batches reject universal gravitation."""
    from dem_mbd_cuda_amd import _abi

    scene, scenes = _small_batch()
    lib = _abi.Library(LIB, "dem_")
    b = scene.SceneBatch(scenes)
    bad = scene.SceneBatch(scenes)
    bad.particle_start = bad.particle_start.copy()
    bad.particle_start[-1] -= 1
    s = lib.solver()
    with pytest.raises(_abi.DemError, match="scene row/wall ranges do not cover"):
        s.initialize_batch(bad)
    empty = scene.SceneBatch(scenes)
    empty.particle_start = empty.particle_start.copy()
    empty.particle_start[1] = empty.particle_start[0]
    with pytest.raises(_abi.DemError, match="scene 0 has an empty or reversed row range"):
        s.initialize_batch(empty)
    uni = scene.SceneBatch(scenes)
    uni.combined.universal_gravity = 1
    with pytest.raises(_abi.DemError, match="cannot use universal gravitation"):
        s.initialize_batch(uni)
    s.close()
    with pytest.raises(ValueError):
        scene.SceneBatch([scenes[0], scene.Scene(**{**scenes[1].__dict__, "dt": 2e-6})])
    assert b.n == sum(sc.n for sc in scenes)
    parts = b.split(np.arange(b.n))
    assert [len(p) for p in parts] == [sc.n for sc in scenes]
