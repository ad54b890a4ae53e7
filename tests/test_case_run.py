"""End to end on the CPU oracle: dem_mbd_cuda_amd.run (the reference host's
case runner: world.par, inputs, step loop, frame files, CSVs, checkpoint)
on the reference's validation cases against every file the reference's own
binary writes for them (tests/golden/ref_case_outputs.npz).  Where the
reference CPU binary's legacy law and the CUDA law coincide (no contacts,
wall contact, particle pair) the files are byte-identical; the body-state
files match up to the six impulse columns that only the coupled CUDA host
appends; the particle-triangle case agrees to 1e-12 relative (the two laws
differ there).  The reference CPU binary writes no contact-history file or
branch CSVs; those are checked for presence and shape."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN
from dem_mbd_cuda_amd import run

REF = np.load(os.path.join(GOLDEN, "ref_case_outputs.npz"))
EXACT = ["free_fall", "wall_contact", "two_particle"]


def outputs_of(case, oracle, tmp_path):
    dst = str(tmp_path / case)
    shutil.copytree(os.path.join(GOLDEN, "cases", case), dst)
    run.run_case(dst, lib=oracle, log=lambda *a: None)
    base = os.path.join(dst, "OutputFile")
    got = {}
    for root, _, files in os.walk(base):
        for f in files:
            got[os.path.relpath(os.path.join(root, f), base)] = os.path.join(root, f)
    return got


def ref_files(case):
    return {k.split("/", 1)[1]: str(REF[k]) for k in REF.files if k.startswith(case + "/")}


@pytest.mark.parametrize("case", EXACT)
def test_byte_identical(case, oracle, tmp_path):
    got = outputs_of(case, oracle, tmp_path)
    ref = ref_files(case)
    assert ref and set(ref) <= set(got)
    for rel, text in ref.items():
        assert open(got[rel]).read() == text, rel
    assert "contact_history.bt" in got


def test_body_states_impulse_columns(oracle, tmp_path):
    got = outputs_of("body_states", oracle, tmp_path)
    ref = ref_files("body_states")
    for rel, text in ref.items():
        mine = open(got[rel]).read()
        if rel.startswith("state_bodys/"):
            for a, b in zip(mine.splitlines(), text.splitlines()):
                assert a.startswith(b) and len(a) == len(b) + 6 * 26, rel
        else:
            assert mine == text, rel
    # one row per body for the initial evaluation and for every output frame
    frames = len([r for r in got if r.startswith("state_bodys/")])
    nb = len(open(got["state_bodys/At.00000000.bt"]).read().splitlines())
    for csv, schema in (("native_body_contact_branches.csv", "native_pt_branch_v1"),
                        ("native_body_contact_microstate.csv", "native_pt_microstate_v1")):
        lines = open(got[csv]).read().splitlines()
        assert lines[0].startswith("schema_version,end_step,end_time_s,body_id,window_native_steps")
        assert len(lines) == 1 + nb * frames
        assert all(ln.startswith(schema + ",") for ln in lines[1:])


def test_triangle_contact_numerically(oracle, tmp_path):
    got = outputs_of("triangle_contact", oracle, tmp_path)
    for rel, text in ref_files("triangle_contact").items():
        a = np.array(open(got[rel]).read().split(), float)
        b = np.array(text.split(), float)
        if rel.startswith("state_bodys/"):
            a = np.array([ln.split()[:20] for ln in open(got[rel]).read().splitlines()], float).ravel()
        assert a.shape == b.shape, rel
        assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1e-3) + 1e-300), rel
