"""Builds the native libraries in-tree (they travel with the source tree).

  libdemhip.so    HIP kernels + C-ABI (include/dem_hip.h), gfx950
  libdemscene.so  seeded sphere-packing builder used by scenes/benchmarks
  libdemio.so     the reference's per-frame output files (ph, F, FB, At)
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# FP64 throughout, no FMA contraction (the law is evaluated operation by
# operation as written); packed-FP32 instructions disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + res.stdout)
    return res.stdout


def build_demhip():
    src = os.path.join(HERE, "csrc", "dem_hip.hip")
    out = os.path.join(HERE, "libdemhip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_scene():
    src = os.path.join(HERE, "csrc", "scene_rsa.c")
    out = os.path.join(HERE, "libdemscene.so")
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", out, src, "-lm"])
    return out


def build_io():
    src = os.path.join(HERE, "csrc", "dem_io.c")
    out = os.path.join(HERE, "libdemio.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-o", out, src])
    return out


def build_all():
    return [build_demhip(), build_scene(), build_io()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
