"""The case runner on the HIP backend: every output file (frames, branch
CSVs, contact-history checkpoint) of the reference's validation cases is
byte-identical to the oracle's, and -- where oracle/_ref/dem_mbd_hip is built
(the reference's own host compiled in CUDA mode, linked to the adapter and
libdemhip.so) -- to what the reference's own host writes through the same
backend."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT
from dem_mbd_cuda_amd import run

pytestmark = pytest.mark.gpu

CASES = ["free_fall", "wall_contact", "two_particle", "triangle_contact", "body_states"]
REF_HOST = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_hip")


def tree(case_dir):
    base = os.path.join(case_dir, "OutputFile")
    out = {}
    for root, _, files in os.walk(base):
        for f in files:
            p = os.path.join(root, f)
            out[os.path.relpath(p, base)] = open(p, "rb").read()
    return out


def fresh(case, tmp_path, tag):
    dst = str(tmp_path / f"{case}_{tag}")
    shutil.copytree(os.path.join(GOLDEN, "cases", case), dst)
    return dst


@pytest.mark.parametrize("case", CASES)
def test_hip_outputs_equal_oracle(case, hip, oracle, tmp_path):
    g, o = fresh(case, tmp_path, "hip"), fresh(case, tmp_path, "ora")
    run.run_case(g, lib=hip, log=lambda *a: None)
    run.run_case(o, lib=oracle, log=lambda *a: None)
    tg, to = tree(g), tree(o)
    assert sorted(tg) == sorted(to) and "contact_history.bt" in tg
    for rel in to:
        assert tg[rel] == to[rel], rel


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host build not present")
@pytest.mark.parametrize("case", CASES)
def test_reference_host_writes_the_same_files(case, hip, tmp_path):
    mine, theirs = fresh(case, tmp_path, "py"), fresh(case, tmp_path, "ref")
    run.run_case(mine, lib=hip, log=lambda *a: None)
    res = subprocess.run([REF_HOST, theirs], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout[-3000:]
    tm, tt = tree(mine), tree(theirs)
    assert sorted(tm) == sorted(tt)
    for rel in tt:
        assert tm[rel] == tt[rel], rel


def clump_case(tmp_path, tag, n=240, steps=40):
    """A walled box of n owners, each a clump of 1-3 spheres, with both
    sidecars named in world.par (principal inertia and clump components),
    gravity, seeded velocities and spins; frames every 10 steps."""
    import numpy as np

    from dem_mbd_cuda_amd import scene

    d = str(tmp_path / f"clumps_{tag}")
    for sub in ("Particles", "Environment", "SettingData"):
        os.makedirs(os.path.join(d, "InputFile", sub))
    sc = scene.box_scene(n=n, seed=31, box=(0.024, 0.024, 0.024), steps=steps, dt=1e-6)
    rng = np.random.RandomState(32)
    cl = scene.clumpify(sc, rng.randint(1, 4, n), seed=33, shrink=0.8, spread=0.35)
    quat = scene.random_quaternions(n, 34)
    vel = rng.uniform(-5.0, 5.0, (n, 3))
    ang = rng.uniform(-300.0, 300.0, (n, 3))
    inp = os.path.join(d, "InputFile")
    with open(os.path.join(inp, "Particles", "Particles.bt"), "w") as f:
        for i in range(n):
            vals = [sc.mass[i], sc.radius[i], *sc.pos[i], *vel[i], *quat[i], *ang[i]]
            f.write(f"{int(sc.pid[i])} 0 " + " ".join(f"{v:.17g}" for v in vals) + "\n")
    with open(os.path.join(inp, "Particles", "ParticleInertia.bt"), "w") as f:
        for i in range(n):
            base = 0.4 * sc.mass[i] * sc.radius[i] ** 2
            f.write(f"{int(sc.pid[i])} {base:.17g} {1.1 * base:.17g} {1.2 * base:.17g}\n")
    with open(os.path.join(inp, "Particles", "ClumpComponents.bt"), "w") as f:
        for c in range(len(cl.comp_id)):
            b = cl.comp_pos_body[c]
            f.write(f"{int(cl.comp_id[c])} {int(sc.pid[cl.comp_owner[c]])} {cl.comp_radius[c]:.17g} "
                    f"{b[0]:.17g} {b[1]:.17g} {b[2]:.17g}\n")
    with open(os.path.join(inp, "Environment", "Wall.bt"), "w") as f:
        f.write(f"{len(sc.wall_origin)}\n")
        for o, nrm in zip(sc.wall_origin, sc.wall_normal):
            f.write(" ".join(f"{v:.17g}" for v in (*o, *nrm)) + "\n")
    m = sc.pp
    par = {"ODE_Option": 1, "ODE_StepSize": "1e-6", "ODE_EndStep": steps, "ODE_StartStep": 0,
           "ODE_OutputInterval": 10, "MECH_mu": m["mu"], "MECH_mu_T": m["muT"], "MECH_mu_R": m["muR"],
           "MECH_epsN": m["epsN"], "MECH_epsS": m["epsS"], "MECH_cohesive": 0, "MECH_beta": m["beta"],
           "MECH_kN": repr(float(m["kN"])), "CONTROL_GRAVITY_FLAG": 1, "CONTROL_GRAVITY_X": 0,
           "CONTROL_GRAVITY_Y": 0, "CONTROL_GRAVITY_Z": -9.81, "CONTROL_WALL_FLAG": 1,
           "CONTROL_WALL_FILE": "Data\\DATA\\InputFile\\Environment\\Wall.bt", "CONTROL_MULTIBODY_FLAG": 0,
           "CONTROL_UNIVERSAL_GRAVITATION_FLAG": 0,
           "CONTROL_PARTICLES_FILE": "Data\\DATA\\InputFile\\Particles\\Particles.bt",
           "CONTROL_PARTICLE_INERTIA_FILE": "Data\\DATA\\InputFile\\Particles\\ParticleInertia.bt",
           "CONTROL_CLUMP_COMPONENTS_FILE": "Data\\DATA\\InputFile\\Particles\\ClumpComponents.bt",
           "CONTROL_OUTPUT_PARTICLE_STATE": 1, "CONTROL_OUTPUT_PARTICLE_FORCE": 1,
           "CONTROL_OUTPUT_BODY_STATE": 0, "CONTROL_OUTPUT_BODY_FORCE": 0, "CONTROL_OUTPUT_SUMMARY": 0}
    with open(os.path.join(inp, "SettingData", "world.par"), "w") as f:
        for k, v in par.items():
            f.write(f"{k} = {v}\n")
    return d


def test_clump_case_hip_equals_oracle(hip, oracle, tmp_path):
    g, o = clump_case(tmp_path, "hip"), clump_case(tmp_path, "ora")
    run.run_case(g, lib=hip, log=lambda *a: None)
    run.run_case(o, lib=oracle, log=lambda *a: None)
    tg, to = tree(g), tree(o)
    assert sorted(tg) == sorted(to) and "contact_history.bt" in tg
    for rel in to:
        assert tg[rel] == to[rel], rel


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host build not present")
def test_clump_case_reference_host_writes_the_same_files(hip, tmp_path):
    """The reference's own host reads the two sidecars itself
    (particle.cpp:108-278) and steps the clumps through the adapter."""
    mine, theirs = clump_case(tmp_path, "py"), clump_case(tmp_path, "ref")
    run.run_case(mine, lib=hip, log=lambda *a: None)
    res = subprocess.run([REF_HOST, theirs], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout[-3000:]
    tm, tt = tree(mine), tree(theirs)
    assert sorted(tm) == sorted(tt)
    for rel in tt:
        assert tm[rel] == tt[rel], rel


def test_prescribed_motion_case(hip, oracle, tmp_path):
    """The triangle validation case with its body prescribed and a schedule
    of three keys (CONTROL_PRESCRIBED_MOTION_FILE): every file the case
    runner writes on the HIP backend equals the oracle's and, where built,
    the reference's own host's through the same backend
    (cuda_simulation_integrator.cpp:172-243, 288-294, 372-373)."""
    from test_prescribed_motion import prescribed_case

    text = "0 0.5 0 0 0 0 1.5\n7 0 -0.25 0 0 0 0\n13 0.1 0.1 0 0 0 -2\n"
    g, _ = prescribed_case(tmp_path, "hip", text)
    o, _ = prescribed_case(tmp_path, "ora", text)
    run.run_case(g, lib=hip, log=lambda *a: None)
    run.run_case(o, lib=oracle, log=lambda *a: None)
    tg, to = tree(g), tree(o)
    assert sorted(tg) == sorted(to) and os.path.join("state_bodys", "At.00000000.bt") in tg
    for rel in to:
        assert tg[rel] == to[rel], rel
    if os.path.exists(REF_HOST):
        r, _ = prescribed_case(tmp_path, "ref", text)
        res = subprocess.run([REF_HOST, r], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=300)
        assert res.returncode == 0, res.stdout[-3000:]
        tr = tree(r)
        assert sorted(tr) == sorted(tg)
        for rel in tr:
            assert tg[rel] == tr[rel], rel
