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
