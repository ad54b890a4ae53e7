"""Writes the golden fixtures under tests/golden/ from the reference's own
validation inputs (data files, not source) and from the reference's CPU
binary built in oracle/_ref/.  Run in the build container, where the
reference tree exists:

    python tests/golden/make_fixtures.py

Outputs (committed):
  case_<name>.npz     the five analytic validation cases (validation/cases/*)
  data1_mbd.npz       DATA1 inputs (validation/ctest_data1_mbd), 1,704
                      spheres, 5 walls, 2 bodies x 1,416 triangles
  data1_pure.npz      the same without bodies (validation/ctest_data1_pure_dem)
  data1_zero_vel_ref_cpu.npz  the initial (frame 0) outputs of the same binary
                      on DATA1 with all particle velocities set to zero
  ref_output_text.npz the text of the reference binary's frame files on DATA1
                      (coupled): ph, F, FB, At of frames 0 and 5, the first 40
                      lines of each (byte-level pin of the output writers)
  cases/<case>/InputFile/...  the input files of the reference's validation
                      cases (validation/cases), as the reference holds them
  ref_case_outputs.npz every OutputFile/ file the reference binary writes for
                      those cases (text), keyed "<case>/<path under OutputFile>"
  law_pins_ref_cpu.npz the numbers in every frame file the reference binary
                      writes for the contact-law and integrator cases of
                      law_cases.py, keyed "<case>/<path under OutputFile>"
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

from dem_mbd_cuda_amd import scene as S  # noqa: E402

CASES = ["free_fall", "wall_contact", "two_particle", "triangle_contact", "body_states"]


def run_reference_cpu(case_dir, work, zero_velocity=False, tag="", all_outputs=True):
    """Runs oracle/_ref/dem_mbd_cpu on a copy of case_dir, with all outputs on
    unless all_outputs is False (optionally with every particle velocity and
    angular velocity zeroed)."""
    dst = os.path.join(work, os.path.basename(case_dir) + tag)
    shutil.copytree(case_dir, dst)
    if zero_velocity:
        pf = os.path.join(dst, "InputFile", "Particles", "Particles.bt")
        rows = np.loadtxt(pf, ndmin=2)
        rows[:, 7:10] = 0.0
        rows[:, 14:17] = 0.0
        with open(pf, "w") as f:
            for r in rows:
                f.write("%d %d " % (int(r[0]), int(r[1])) + " ".join("%.17e" % v for v in r[2:]) + "\n")
    par = os.path.join(dst, "InputFile", "SettingData", "world.par")
    lines = open(par, encoding="utf-8", errors="replace").read().splitlines()
    out = []
    for ln in lines:
        key = ln.split("=")[0].strip()
        if all_outputs and key.startswith("CONTROL_OUTPUT_") and key != "CONTROL_OUTPUT_SUMMARY":
            ln = f"{key} = 1"
        out.append(ln)
    open(par, "w").write("\n".join(out) + "\n")
    exe = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_cpu")
    res = subprocess.run([exe, dst], cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-4000:])
    return dst


def read_outputs(case_out):
    d = {}
    for sub, pref in (("state_particles", "ph"), ("force_particles", "F"), ("state_bodys", "At")):
        for f in sorted(glob.glob(os.path.join(case_out, "OutputFile", sub, pref + ".*.bt"))):
            idx = int(os.path.basename(f).split(".")[1])
            d[f"{pref}_{idx}"] = np.loadtxt(f, ndmin=2)
    return d


def main():
    for name in CASES:
        sc = S.read_case(os.path.join(REF, "validation", "cases", name))
        sc.save(os.path.join(HERE, f"case_{name}.npz"))
    for tag, d in (("mbd", "ctest_data1_mbd"), ("pure", "ctest_data1_pure_dem")):
        sc = S.read_case(os.path.join(REF, "validation", d))
        sc.save(os.path.join(HERE, f"data1_{tag}.npz"))
    zref = {}
    with tempfile.TemporaryDirectory() as work:
        for tag, d in (("mbd", "ctest_data1_mbd"), ("pure", "ctest_data1_pure_dem")):
            out = run_reference_cpu(os.path.join(REF, "validation", d), work, zero_velocity=True, tag="_zv")
            for k, v in read_outputs(out).items():
                if k.endswith("_0"):
                    zref[f"{tag}__{k}"] = v
    np.savez_compressed(os.path.join(HERE, "data1_zero_vel_ref_cpu.npz"), **zref)
    text = {}
    with tempfile.TemporaryDirectory() as work:
        out = run_reference_cpu(os.path.join(REF, "validation", "ctest_data1_mbd"), work)
        for sub, name in (("state_particles", "ph.%08d.bt"), ("force_particles", "F.%08d.bt"),
                          ("force_bodys", "FB.%05d.bt"), ("state_bodys", "At.%08d.bt")):
            for idx in (0, 5):
                f = os.path.join(out, "OutputFile", sub, name % idx)
                lines = open(f).read().splitlines(keepends=True)[:40]
                text[name.split(".")[0] + f"_{idx}"] = np.array("".join(lines))
    np.savez_compressed(os.path.join(HERE, "ref_output_text.npz"), **text)
    outs = {}
    with tempfile.TemporaryDirectory() as work:
        for name in CASES:
            src = os.path.join(REF, "validation", "cases", name)
            dst = os.path.join(HERE, "cases", name)
            if os.path.exists(dst):
                shutil.rmtree(dst)
            shutil.copytree(os.path.join(src, "InputFile"), os.path.join(dst, "InputFile"))
            for root, dirs, files in os.walk(dst):
                for f in files:
                    os.chmod(os.path.join(root, f), 0o644)
                os.chmod(root, 0o755)
            out = run_reference_cpu(src, work, all_outputs=False)
            base = os.path.join(out, "OutputFile")
            for root, _, files in os.walk(base):
                for f in sorted(files):
                    rel = os.path.relpath(os.path.join(root, f), base)
                    outs[f"{name}/{rel}"] = np.array(open(os.path.join(root, f)).read())
    np.savez_compressed(os.path.join(HERE, "ref_case_outputs.npz"), **outs)
    law_pin_fixtures()
    print("fixtures written to", HERE)


def law_pin_fixtures():
    import law_cases

    exe = os.path.join(ROOT, "oracle", "_ref", "dem_mbd_cpu")
    outs = {}
    with tempfile.TemporaryDirectory() as work:
        for name in law_cases.CASES:
            d = law_cases.write_case(name, os.path.join(work, name))
            res = subprocess.run([exe, d], cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                 text=True, timeout=600)
            if res.returncode != 0:
                raise RuntimeError(res.stdout[-4000:])
            base = os.path.join(d, "OutputFile")
            for f in sorted(glob.glob(os.path.join(base, "*", "*.bt"))):
                outs[f"{name}/{os.path.relpath(f, base)}"] = np.loadtxt(f, ndmin=2)
    np.savez_compressed(os.path.join(HERE, "law_pins_ref_cpu.npz"), **outs)


if __name__ == "__main__":
    if sys.argv[1:] == ["law_pins"]:
        law_pin_fixtures()
    else:
        main()
