"""Differentiable scene observables for torch.autograd (include/dem_hip.h,
dem_diff_forward / dem_diff_backward).

`BedObservables.apply(params, scenes, steps, checkpoint_interval, plate_wall)`
takes per-scene contact parameters params[s] = (mu, restitution, E) as a
float64 tensor, applies them to both the particle-particle and the
particle-wall law of every scene (epsN = epsS = restitution, kN = E * the
scene's mean radius, as in the sweep scenes), runs the batch on the GPU and
returns [s] = (mean bed height after the last step, time-averaged plate load).
Its backward runs the hand-written adjoint (checkpointed every
`checkpoint_interval` steps, recomputed in between) and maps the law
derivatives back onto (mu, restitution, E).
"""
from __future__ import annotations

import numpy as np
import torch

from . import library
from .scene import SceneBatch


def apply_params(scenes, params):
    """Copies of the scenes with (mu, restitution, E) per scene applied."""
    out = []
    for sc, (mu, e, E) in zip(scenes, np.asarray(params, dtype=np.float64)):
        s = sc.copy()
        kN = float(E) * float(np.mean(sc.radius))
        for key in ("pp", "pw"):
            m = dict(getattr(sc, key))
            m.update(mu=float(mu), epsN=float(e), epsS=float(e), kN=kN)
            setattr(s, key, m)
        out.append(s)
    return out


class BedObservables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, scenes, steps, checkpoint_interval, plate_wall):
        p = params.detach().cpu().numpy().astype(np.float64)
        batch = SceneBatch(apply_params(scenes, p))
        solver = library().solver()
        solver.initialize_batch(batch)
        h, f = solver.diff_forward(steps, checkpoint_interval, plate_wall, scenes=len(scenes))
        ctx.solver = solver
        ctx.rbar = np.array([float(np.mean(sc.radius)) for sc in scenes])
        return torch.from_numpy(np.stack([h, f], axis=1))

    @staticmethod
    def backward(ctx, grad):
        g = grad.detach().cpu().numpy().astype(np.float64)
        d = ctx.solver.diff_backward(g[:, 0], g[:, 1])
        d_mu = d[:, 0] + d[:, 4]
        d_e = d[:, 1] + d[:, 2] + d[:, 5] + d[:, 6]
        d_E = (d[:, 3] + d[:, 7]) * ctx.rbar
        out = torch.from_numpy(np.stack([d_mu, d_e, d_E], axis=1))
        return out, None, None, None, None


def observables(scenes, params, steps, checkpoint_interval=100, plate_wall=4):
    """Forward only (no gradient bookkeeping): [s] = (bed height, plate load)."""
    batch = SceneBatch(apply_params(scenes, params))
    solver = library().solver()
    solver.initialize_batch(batch)
    h, f = solver.diff_forward(steps, checkpoint_interval, plate_wall, scenes=len(scenes))
    solver.close()
    return np.stack([h, f], axis=1)
