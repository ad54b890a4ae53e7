"""The hand-written reverse-mode derivatives of the particle-particle and
particle-wall laws (dem_adjoint.hpp) against central finite differences of
the forward laws, both evaluated on the CPU from libdemhip.so's host copies
of the law arithmetic.  Random contacts in each branch: elastic (no yield),
sliding (Coulomb cap active), rotating normals (transported history), the
first evaluation (dt = 0) and frozen partners (status -1)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "dem_mbd_cuda_amd", "libdemhip.so")


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    for name in ("dem_check_pp_law", "dem_check_pw_law"):
        getattr(L, name).argtypes = [dp, dp]
    for name in ("dem_check_pp_law_vjp", "dem_check_pw_law_vjp"):
        getattr(L, name).argtypes = [dp, dp, dp]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def call(L, name, x, nout):
    x = np.ascontiguousarray(x, float)
    out = np.zeros(nout)
    assert getattr(L, name)(_p(x), _p(out)) == 0
    return out


def vjp(L, name, x, seeds, ngrad):
    x = np.ascontiguousarray(x, float)
    s = np.ascontiguousarray(seeds, float)
    g = np.zeros(ngrad)
    assert getattr(L, name)(_p(x), _p(s), _p(g)) == 0
    return g


def body(rng, pos, r, status=0):
    return np.concatenate([pos, rng.normal(0, 0.3, 3), rng.normal(0, 0.3, 3), rng.normal(0, 40, 3),
                           rng.normal(0, 40, 3), [0, 0, 0], [2500 * 4 / 3 * np.pi * r ** 3, r, status]])


def mech(rng, dt):
    # epsS mu epsN cohesion beta muT muR kN dt
    return np.array([rng.uniform(0.2, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.2, 0.9), 0.0, 0.5, 0.1, 0.1,
                     10 ** rng.uniform(3.5, 5.0), dt])


# inputs checked: body pos/vel/vh/w/wh (15 per body), slot values (9),
# material mu, epsN, epsS, kN through their gradient entries
def fd_check(L, name, x, nout, body_slices, prev_at, mech_at, ngrad, grad_prev_at, grad_mat_at, rng,
             rtol=2e-6):
    seeds = rng.normal(size=nout)
    g = vjp(L, name + "_vjp", x, seeds, ngrad)
    f = lambda z: call(L, name, z, nout) @ seeds
    checks = []
    for bi, (start, gstart) in enumerate(body_slices):
        for k in range(15):
            checks.append((start + k, g[gstart + k]))
    for k in range(9):
        checks.append((prev_at + k, g[grad_prev_at + k]))
    # mu
    checks.append((mech_at + 1, g[grad_mat_at + 0]))
    # epsN, epsS: gradient reported for log(eps) -> d/deps = g / eps
    checks.append((mech_at + 2, g[grad_mat_at + 1] / x[mech_at + 2]))
    checks.append((mech_at + 0, g[grad_mat_at + 2] / x[mech_at + 0]))
    checks.append((mech_at + 7, g[grad_mat_at + 3]))
    scale = max(1e-300, max(abs(v) for _, v in checks))
    for idx, analytic in checks:
        h = 1e-6 * max(abs(x[idx]), 1e-6)
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (f(xp) - f(xm)) / (2 * h)
        assert abs(fd - analytic) <= rtol * scale + 1e-7 * abs(fd), (name, idx, fd, analytic, scale)


@pytest.mark.parametrize("case", ["elastic", "sliding", "first", "frozen", "big_history"])
@pytest.mark.parametrize("seed", range(4))
def test_pp_law_vjp(lib, case, seed):
    rng = np.random.RandomState(100 * seed + hash(case) % 97)
    ri, rj = rng.uniform(1e-3, 1.5e-3, 2)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    dist = (ri + rj) * (1 - rng.uniform(0.002, 0.02))
    pi = rng.normal(0, 1e-2, 3)
    I = body(rng, pi, ri)
    J = body(rng, pi + d * dist, rj, status=-1 if case == "frozen" else 0)
    m = mech(rng, 0.0 if case == "first" else 1e-5)
    prev = rng.normal(0, 1e-6, 9)
    if case == "sliding":
        m[1] = 0.05  # small mu: the Coulomb cap binds
        prev = rng.normal(0, 1e-4, 9)
    if case == "big_history":
        prev = rng.normal(0, 1e-3, 9)
    x = np.concatenate([I, J, m, prev])
    fd_check(lib, "dem_check_pp_law", x, 15, [(0, 0), (21, 15)], 51, 42, 43, 30, 39, rng)


@pytest.mark.parametrize("case", ["elastic", "sliding", "first"])
@pytest.mark.parametrize("seed", range(4))
def test_pw_law_vjp(lib, case, seed):
    rng = np.random.RandomState(7 + 31 * seed + hash(case) % 89)
    r = rng.uniform(1e-3, 1.5e-3)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    origin = rng.normal(0, 1e-2, 3)
    pos = origin + n * r * (1 - rng.uniform(0.002, 0.02)) + np.cross(n, rng.normal(0, 1e-3, 3))
    I = body(rng, pos, r)
    m = mech(rng, 0.0 if case == "first" else 1e-5)
    prev = rng.normal(0, 1e-6, 9)
    if case == "sliding":
        m[1] = 0.05
        prev = rng.normal(0, 1e-4, 9)
    x = np.concatenate([I, n, origin, m, prev])
    fd_check(lib, "dem_check_pw_law", x, 15, [(0, 0)], 36, 27, 28, 15, 24, rng)
