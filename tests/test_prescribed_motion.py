"""The prescribed-motion schedule (CONTROL_PRESCRIBED_MOTION_FILE; reference
src/cuda_simulation_integrator.cpp:172-243, 288-294, 372-373) on the CPU.

* load_prescribed_motion keeps the reference's row rules and messages: '#'
  comments, commas as separators, rows whose first field is not an integer
  skipped, six values per row, strictly increasing steps inside
  [start, end), the first at the start step.
* Where oracle/_ref/dem_mbd_hip is built (the reference's own host in CUDA
  mode), every rejected file is also run through it: the host parses the
  schedule before it touches a device, so it reports the same message here,
  from the reference's own code.
* run_coupled applies the first key before the first force evaluation (the
  frame before step 0 carries it) and refuses a schedule unless exactly one
  body, prescribed, exists, before any evaluation.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from dem_mbd_cuda_amd import host, run

REF_HOST = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_hip")

BAD = [
    ("missing", None, "cannot open prescribed motion file: {path}"),
    ("short row", "0 1 2 3 4 5\n", "invalid prescribed motion row 1"),
    ("non-number", "# header\n0 1 2 3 x 5 6\n", "invalid prescribed motion row 2"),
    ("extra column", "0,0,0,0,0,0,1,9\n", "extra prescribed motion column at row 1"),
    ("repeated step", "0 0 0 0 0 0 1\n3 0 0 0 0 0 1\n3 0 0 0 0 0 2\n",
     "prescribed motion steps must increase within the integration range"),
    ("decreasing step", "0 0 0 0 0 0 1\n5 0 0 0 0 0 1\n4 0 0 0 0 0 2\n",
     "prescribed motion steps must increase within the integration range"),
    ("step at the end", "0 0 0 0 0 0 1\n20 0 0 0 0 0 1\n",
     "prescribed motion steps must increase within the integration range"),
    ("negative step", "-1 0 0 0 0 0 1\n", "prescribed motion steps must increase within the integration range"),
    ("late start", "2 0 0 0 0 0 1\n", "prescribed motion must begin at ODE_StartStep"),
    ("empty", "# nothing\n\n", "prescribed motion must begin at ODE_StartStep"),
]


def write(tmp_path, name, text):
    p = str(tmp_path / name)
    if text is not None:
        with open(p, "w") as f:
            f.write(text)
    return p


@pytest.mark.parametrize("name,text,msg", BAD, ids=[b[0] for b in BAD])
def test_rejected_schedules(tmp_path, name, text, msg):
    path = write(tmp_path, "motion.csv", text)
    with pytest.raises(RuntimeError) as e:
        host.load_prescribed_motion(path, 0, 20)
    assert str(e.value) == msg.format(path=path)


def test_accepted_schedule_rows(tmp_path):
    text = ("step,vx,vy,vz,wx,wy,wz   # a header row is skipped\n"
            "0, 0.5, 0, 0, 0, 0, 1.5\n"
            "\n"
            "   # comment only\n"
            "7 -1e-3 2 3 4 5 6  # trailing comment\n"
            "13\t0\t0\t0\t0\t0\t-2\n")
    keys = host.load_prescribed_motion(write(tmp_path, "motion.csv", text), 0, 20)
    assert [k[0] for k in keys] == [0, 7, 13]
    assert np.array_equal(keys[1][1], [-1e-3, 2, 3]) and np.array_equal(keys[1][2], [4, 5, 6])
    assert np.array_equal(keys[2][2], [0, 0, -2])
    assert host.load_prescribed_motion("", 0, 20) == []
    # the first key at a later start step
    assert host.load_prescribed_motion(write(tmp_path, "m2.csv", "5 0 0 0 0 0 1\n"), 5, 9)[0][0] == 5


def prescribed_case(tmp_path, tag, motion_text, state=2, end=20):
    """The reference's triangle validation case with its body prescribed
    (bodies.csv state 2), 20 steps, frames every 5, and a schedule file
    named by absolute path in world.par."""
    d = str(tmp_path / f"pm_{tag}")
    shutil.copytree(os.path.join(GOLDEN, "cases", "triangle_contact"), d)
    csv = os.path.join(d, "InputFile", "BodySet", "bodies.csv")
    lines = open(csv).read().splitlines()
    lines[2] = lines[2].rsplit(",", 1)[0] + f",{state}"
    open(csv, "w").write("\n".join(lines) + "\n")
    motion = write(tmp_path, f"motion_{tag}.csv", motion_text)
    par = os.path.join(d, "InputFile", "SettingData", "world.par")
    text = open(par).read().replace("ODE_EndStep = 1\n", f"ODE_EndStep = {end}\n")
    text = text.replace("ODE_OutputInterval = 1\n", "ODE_OutputInterval = 5\n")
    open(par, "w").write(text + f"CONTROL_PRESCRIBED_MOTION_FILE = {motion}\n")
    return d, motion


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host build not present")
@pytest.mark.parametrize("name,text,msg", [b for b in BAD if b[0] != "missing"] + [BAD[0]],
                         ids=[b[0] for b in BAD if b[0] != "missing"] + ["missing"])
def test_reference_host_rejects_the_same_files(tmp_path, name, text, msg, oracle):
    d, motion = prescribed_case(tmp_path, "ref", text)
    res = subprocess.run([REF_HOST, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert res.returncode == 1
    want = msg.format(path=motion)
    assert f"Reason / 原因: {want}\n" in res.stdout, res.stdout[-800:]
    # the case runner stops with the same message before stepping
    mine, mine_motion = prescribed_case(tmp_path, "mine", text)
    with pytest.raises(RuntimeError) as e:
        run.run_case(mine, lib=oracle, log=lambda *a: None)
    assert str(e.value) == msg.format(path=mine_motion)


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host build not present")
def test_reference_host_requires_one_prescribed_body(tmp_path, oracle):
    """A dynamic body under a valid schedule: both hosts stop before the
    first evaluation."""
    good = "0 0 0 0 0 0 1\n"
    d, _ = prescribed_case(tmp_path, "ref", good, state=0)
    res = subprocess.run([REF_HOST, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert res.returncode == 1
    msg = "prescribed motion schedule currently requires one prescribed body"
    assert f"Reason / 原因: {msg}\n" in res.stdout, res.stdout[-800:]
    mine, _ = prescribed_case(tmp_path, "mine", good, state=0)
    with pytest.raises(RuntimeError) as e:
        run.run_case(mine, lib=oracle, log=lambda *a: None)
    assert str(e.value) == msg
    assert not os.path.exists(os.path.join(mine, "OutputFile", "state_bodys", "At.00000000.bt"))


def test_first_key_applies_before_the_first_evaluation(tmp_path, oracle):
    """Keys at steps 0, 7, 13 with velocities unlike bodies.csv (which holds
    zero): the frame before step 0 carries key 0; the body then moves with
    key 0 for steps 0-6, key 1 for 7-12 and key 2 after, so its centre is
    the sum of those displacements (to rounding), and its orientation turns
    by the integrated angular velocity."""
    from dem_mbd_cuda_amd import scene

    d, motion = prescribed_case(tmp_path, "run", "0 0.5 0 0 0 0 1.5\n7 0 -0.25 0 0 0 0\n13 0.1 0.1 0 0 0 -2\n")
    sc = scene.read_case(d)
    keys = host.load_prescribed_motion(motion, 0, sc.steps)
    s = oracle.solver()
    s.initialize(sc)
    frames = {}
    host.run_coupled(s, sc, output_interval=5, prescribed=keys,
                     on_frame=lambda step, t, extra: frames.__setitem__(step + 1, extra))
    s.close()
    assert np.array_equal(frames[0]["bodies"]["vel"][0], [0.5, 0.0, 0.0])
    assert np.array_equal(frames[0]["bodies"]["angvel"][0], [0.0, 0.0, 1.5])
    dt = sc.dt
    want = 7 * dt * np.array([0.5, 0, 0]) + 6 * dt * np.array([0, -0.25, 0]) + 7 * dt * np.array([0.1, 0.1, 0])
    assert np.allclose(frames[20]["bodies"]["center"][0], want, rtol=0, atol=1e-15)
    angle = 7 * dt * 1.5 + 7 * dt * (-2.0)
    q = frames[20]["bodies"]["orien"][0]
    assert abs(q[0] - np.cos(angle / 2)) < 1e-12 and abs(q[3] - np.sin(angle / 2)) < 1e-12
    # the frame before step 0 used key 0's velocity in its evaluation: the
    # same case with key 0 = zero gives another frame-0 force on the sphere
    d2, motion2 = prescribed_case(tmp_path, "zero", "0 0 0 0 0 0 0\n7 0 -0.25 0 0 0 0\n")
    sc2 = scene.read_case(d2)
    loads = []
    for c, k in ((sc, keys), (sc2, host.load_prescribed_motion(motion2, 0, sc2.steps))):
        s = oracle.solver()
        s.initialize(c)
        got = {}
        host.run_coupled(s, c, steps=1, prescribed=k,
                         on_frame=lambda step, t, extra: got.setdefault(step + 1, extra["loads"].copy()))
        s.close()
        loads.append(got[0])
    assert not np.array_equal(loads[0], loads[1])
