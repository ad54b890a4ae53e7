"""Shared drivers for the parity tests: run one scene through a dem_hip.h
library (the HIP path or the oracle) with the reference's step loop and
collect every output frame."""
from __future__ import annotations

import numpy as np

from dem_mbd_cuda_amd import host


def run_frames(lib, scene, steps=None, interval=None, debug=False):
    s = lib.solver()
    s.initialize(scene)
    frames = {}

    def cb(step, t, extra):
        fr = dict(state=s.state(), forces=s.forces(), stats=s.stats(), extra=extra,
                  counters=s.run_counters())
        if debug:
            fr["grid"] = s.debug_grid()
            fr["history"] = s.debug_history()
            if s.nb:
                fr["tgrid"] = s.debug_triangle_grid()
        frames[step + 1] = fr

    if len(scene.body_mass):
        host.run_coupled(s, scene, steps=steps, output_interval=interval, on_frame=cb)
    else:
        host.run_pure(s, scene, steps=steps, output_interval=interval, on_frame=cb)
    s.close()
    return frames


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def max_abs(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def assert_close(a, b, rtol, atol, what):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.abs(a - b) > atol + rtol * np.abs(b)
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: {int(bad.sum())} values differ, first at {tuple(idx)}: "
                             f"{a[tuple(idx)]!r} vs {b[tuple(idx)]!r}")
