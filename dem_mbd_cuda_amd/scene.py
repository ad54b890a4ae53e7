"""Scene description: the inputs of one DEM run in the solver's array layout.

A Scene holds what the reference host reads from `world.par`,
`Particles.bt`, `Wall.bt`, `CouplePolyFileName.csv` and the `.bt` meshes
(reference src/simulation.cpp:158-307, src/particle.cpp:18-106,
src/wall.cpp:8-38, src/multibody.cpp:13-145,295-435), as numpy arrays in
the layout of include/dem_hip.h.  It also builds the seeded synthetic
packings used by the benchmarks.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# MECH defaults of the reference (include/simulation.h:48-64).
MECH_DEFAULTS = dict(epsS=1.0, mu=0.5, epsN=1.0, cohesion=1.0, beta=1.0, muT=1.3, muR=1.0, kN=1.0)


def mech(**kw):
    m = dict(MECH_DEFAULTS)
    m.update(kw)
    return m


@dataclasses.dataclass
class Scene:
    pid: np.ndarray
    status: np.ndarray
    mass: np.ndarray
    radius: np.ndarray
    inertia: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    angvel: np.ndarray
    quat: np.ndarray
    comp_id: np.ndarray = None
    comp_owner: np.ndarray = None
    comp_radius: np.ndarray = None
    comp_pos_body: np.ndarray = None
    wall_origin: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3)))
    wall_normal: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3)))
    body_mass: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    body_pose: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 13)))
    body_tri_count: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    tri_local: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 15)))
    body_state: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    body_inertia: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3, 3)))
    pp: dict = dataclasses.field(default_factory=lambda: mech())
    pt: dict = dataclasses.field(default_factory=lambda: mech())
    pw: dict = dataclasses.field(default_factory=lambda: mech())
    dt: float = 1e-6
    gravity: tuple = (0.0, 0.0, 0.0)
    gravity_enabled: int = 0
    universal_gravity: int = 0
    keep_body_particle: int = 0
    keep_body_detail: int = 0
    mesh_scale: float = 0.0
    steps: int = 0
    output_interval: int = 0

    def __post_init__(self):
        n = len(self.pid)
        if self.comp_id is None:
            # default one-sphere topology (reference src/particle.cpp:182-198)
            self.comp_id = np.asarray(self.pid, np.int32).copy()
            self.comp_owner = np.arange(n, dtype=np.int32)
            self.comp_radius = np.asarray(self.radius, np.float64).copy()
            self.comp_pos_body = np.zeros((n, 3))

    @property
    def n(self):
        return len(self.pid)

    def save(self, path):
        d = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        flat = {}
        for k, v in d.items():
            if isinstance(v, dict):
                for kk, vv in v.items():
                    flat[f"{k}__{kk}"] = np.asarray(vv)
            else:
                flat[k] = np.asarray(v)
        np.savez_compressed(path, **flat)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        kw = {}
        mechs = {"pp": {}, "pt": {}, "pw": {}}
        for k in z.files:
            if "__" in k:
                a, b = k.split("__", 1)
                mechs[a][b] = float(z[k])
            else:
                v = z[k]
                kw[k] = v.item() if v.ndim == 0 else v
        kw.update(mechs)
        kw["gravity"] = tuple(np.asarray(kw["gravity"]).tolist())
        return cls(**kw)

    def copy(self):
        return dataclasses.replace(self, **{f.name: (getattr(self, f.name).copy()
                                                      if isinstance(getattr(self, f.name), np.ndarray)
                                                      else getattr(self, f.name))
                                            for f in dataclasses.fields(self)})


def sphere_particles(pos, radius, density=None, mass=None, ids=None, status=None):
    """Owner arrays for solid spheres: I = 0.4 m r^2 (reference particle.cpp:86-90)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    radius = np.asarray(radius, np.float64)
    n = len(radius)
    if mass is None:
        mass = density * (4.0 / 3.0) * np.pi * radius ** 3
    mass = np.asarray(mass, np.float64) * np.ones(n)
    inertia = 0.4 * mass * radius * radius
    return dict(pid=np.arange(n, dtype=np.int32) if ids is None else np.asarray(ids, np.int32),
                status=np.zeros(n, np.int32) if status is None else np.asarray(status, np.int32),
                mass=mass, radius=radius, inertia=np.repeat(inertia[:, None], 3, axis=1),
                pos=pos.copy(), vel=np.zeros((n, 3)), angvel=np.zeros((n, 3)),
                quat=np.tile([1.0, 0.0, 0.0, 0.0], (n, 1)))


def clumpify(scene, counts, seed=0, shrink=0.7, spread=None):
    """Replaces every sphere owner of `scene` by a clump of counts[i] (1..8)
    sphere components: radius shrink * r, centres at spread * r (default
    1 - shrink: the clump stays inside the original sphere) from the owner
    centre along seeded random body-frame directions.  Owner mass and
    inertia are kept; component ids are 8 * owner id + k.  Returns a new
    Scene."""
    spread = 1.0 - shrink if spread is None else spread
    counts = np.broadcast_to(np.asarray(counts, int), (scene.n,))
    if np.any(counts < 1) or np.any(counts > 8):
        raise ValueError("1 to 8 components per owner")
    rng = np.random.RandomState(seed)
    cid, owner, radius, posb = [], [], [], []
    for i in range(scene.n):
        r = float(scene.radius[i])
        d = rng.normal(size=(counts[i], 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        for k in range(counts[i]):
            cid.append(8 * int(scene.pid[i]) + k)
            owner.append(i)
            radius.append(shrink * r)
            posb.append(spread * r * d[k])
    out = scene.copy()
    out.comp_id = np.array(cid, np.int32)
    out.comp_owner = np.array(owner, np.int32)
    out.comp_radius = np.array(radius)
    out.comp_pos_body = np.array(posb).reshape(-1, 3)
    return out


def random_quaternions(n, seed):
    """Seeded uniformly distributed unit quaternions (w, x, y, z)."""
    q = np.random.RandomState(seed).normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


def box_walls(lo, hi):
    """Six planar walls of an axis-aligned box, normals pointing inward."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    origin, normal = [], []
    for a in range(3):
        e = np.zeros(3)
        e[a] = 1.0
        origin.append(lo.copy()); normal.append(e.copy())
        origin.append(hi.copy()); normal.append(-e)
    return np.array(origin), np.array(normal)


_scene_lib = None


def _rsa_lib():
    global _scene_lib
    if _scene_lib is None:
        path = os.path.join(_HERE, "libdemscene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.dem_rsa_box.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_uint64, C.c_int, C.c_double,
                                    C.POINTER(C.c_double)]
        lib.dem_rsa_box.restype = C.c_int
        _scene_lib = lib
    return _scene_lib


def layered_lattice(radius, lo, hi_xy, factor=1.0):
    """Spheres sorted by decreasing radius, laid on square layers from the
    floor up; a layer's spacing is factor * twice its largest radius.
    Returns (radius in placement order, positions)."""
    r = np.sort(np.asarray(radius, float))[::-1].copy()
    n = len(r)
    pos = np.zeros((n, 3))
    wx, wy = hi_xy[0] - lo[0], hi_xy[1] - lo[1]
    z = lo[2]
    prev = 0.0
    i = 0
    while i < n:
        r0 = r[i]
        s = 2.0 * r0 * factor
        nx, ny = max(1, int(wx // s)), max(1, int(wy // s))
        m = min(nx * ny, n - i)
        z = z + (prev + r0) * factor if i else lo[2] + r0
        k = np.arange(m)
        pos[i:i + m, 0] = lo[0] + (wx - nx * s) / 2 + (k % nx + 0.5) * s
        pos[i:i + m, 1] = lo[1] + (wy - ny * s) / 2 + (k // nx + 0.5) * s
        pos[i:i + m, 2] = z
        prev = r0
        i += m
    return r, pos


def relax_packing(radius, pos, lo, hi_xy, max_iter=200, tol=1e-3):
    """Removes overlaps in place (see dem_relax_packing); returns sweeps used."""
    lib = _rsa_lib()
    fn = lib.dem_relax_packing
    dp = C.POINTER(C.c_double)
    fn.argtypes = [C.c_int, dp, dp, dp, C.c_int, C.c_double, dp]
    fn.restype = C.c_int
    radius = np.ascontiguousarray(radius, np.float64)
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray([hi_xy[0], hi_xy[1], 1e30], np.float64)
    it = fn(len(radius), radius.ctypes.data_as(dp), lo.ctypes.data_as(dp), hi.ctypes.data_as(dp),
            int(max_iter), float(tol), pos.ctypes.data_as(dp))
    if it < 0:
        raise RuntimeError("packing relaxation: out of memory")
    return it


def deposit_packing(radius, lo, hi_xy, wall_origin, wall_normal, seed, preload=1e-4, batch=32):
    """Sequential deposition under gravity (see dem_deposit_packing): the
    spheres, in the given order, dropped at seeded (x, y) over the floor
    footprint [lo, hi_xy] and rolled down into a hollow of their supports,
    with the planes (wall_origin, wall_normal) as obstacles.  Resting
    contacts overlap by `preload` of the radius sum.  `batch` spheres are
    computed at a time (in parallel); the result is the same for any batch.
    Returns positions."""
    lib = _rsa_lib()
    fn = lib.dem_deposit_packing
    dp = C.POINTER(C.c_double)
    fn.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp, dp, C.c_uint64, C.c_double, C.c_int, dp]
    fn.restype = C.c_int
    radius = np.ascontiguousarray(radius, np.float64)
    lo2 = np.ascontiguousarray(np.asarray(lo, float)[:2])
    hi2 = np.ascontiguousarray(np.asarray(hi_xy, float)[:2])
    wo = np.ascontiguousarray(wall_origin, np.float64)
    wn = np.ascontiguousarray(wall_normal, np.float64)
    pos = np.zeros((len(radius), 3))
    rc = fn(len(radius), radius.ctypes.data_as(dp), lo2.ctypes.data_as(dp), hi2.ctypes.data_as(dp), len(wo),
            wo.ctypes.data_as(dp), wn.ctypes.data_as(dp), int(seed) & ((1 << 64) - 1), float(preload),
            int(batch), pos.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError("sequential deposition: out of memory")
    return pos


def rsa_positions(radius, lo, hi, seed, max_attempts=2000, gap=0.0):
    radius = np.ascontiguousarray(radius, np.float64)
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    pos = np.zeros((len(radius), 3))
    dp = C.POINTER(C.c_double)
    placed = _rsa_lib().dem_rsa_box(len(radius), radius.ctypes.data_as(dp), lo.ctypes.data_as(dp),
                                    hi.ctypes.data_as(dp), int(seed) & ((1 << 64) - 1),
                                    int(max_attempts), float(gap), pos.ctypes.data_as(dp))
    if placed != len(radius):
        raise RuntimeError(f"random sequential insertion placed only {placed}/{len(radius)} spheres")
    return pos


def box_scene(n=2000, seed=1, box=(0.04, 0.04, 0.08), rmin=1.0e-3, rmax=1.5e-3, density=2500.0,
              E=1.0e7, mu=0.5, restitution=0.5, dt=1.0e-6, steps=2000, muT=0.1, muR=0.1, beta=0.5,
              gravity=(0.0, 0.0, -9.81)):
    """Pure-DEM box: n spheres, radius U[rmin, rmax], seeded RSA, 6 walls.

    The reference's law takes a normal stiffness kN [N/m] (ν is unused); the
    mapping kN = E * mean radius is used, with epsS = epsN = restitution and
    no cohesion.
    """
    rng = np.random.RandomState(seed)
    radius = rng.uniform(rmin, rmax, n)
    lo, hi = np.zeros(3), np.asarray(box, float)
    pos = rsa_positions(radius, lo, hi, seed=seed * 7919 + 17)
    kN = E * 0.5 * (rmin + rmax)
    m = mech(epsS=restitution, mu=mu, epsN=restitution, cohesion=0.0, beta=beta, muT=muT, muR=muR,
             kN=kN)
    wo, wn = box_walls(lo, hi)
    return Scene(**sphere_particles(pos, radius, density), wall_origin=wo, wall_normal=wn,
                 pp=m, pt=dict(m), pw=dict(m), dt=dt, gravity=tuple(gravity), gravity_enabled=1,
                 steps=steps)


def hopper_walls(half_width, angle_deg):
    """Floor z=0 plus four planes leaning outward by angle_deg from vertical."""
    t = np.tan(np.radians(angle_deg))
    origin = [[0.0, 0.0, 0.0]]
    normal = [[0.0, 0.0, 1.0]]
    for a in (0, 1):
        for sgn in (1.0, -1.0):
            o = np.zeros(3)
            o[a] = -sgn * half_width
            nrm = np.zeros(3)
            nrm[a] = sgn
            nrm[2] = t
            origin.append(o)
            normal.append(nrm / np.linalg.norm(nrm))
    return np.array(origin), np.array(normal)


def hopper_scene(n=1_000_000, seed=2, width=0.2, fill=0.36, median=1.0e-3, sigma=0.2,
                 rclip=(0.5e-3, 2.0e-3), density=2500.0, E=1.0e7, mu=0.5, restitution=0.5,
                 dt=5.0e-7, steps=5000, angle_deg=10.0, muT=0.1, muR=0.1, beta=0.5,
                 mesh_scale=0.0, gravity=(0.0, 0.0, -9.81), packing="rsa"):
    """Pure-DEM hopper: lognormal radii, floor + 4 inclined walls, gravity.

    packing="rsa": seeded random sequential insertion into the hopper's
    footprint column at volume fraction `fill` (largest radii first; loose,
    no initial contacts).  packing="settled": the spheres deposited one by
    one under gravity, each rolled into a hollow of the bed (sequential
    deposition, see deposit_packing) with resting contacts preloaded by 1e-4
    of the radius sum (0.2 um, the order of the weight-induced overlap of
    the stiffness kN = E * median at mid-depth): a bed at rest with about
    5.5 contacts per sphere.  packing="dense": size-sorted lattice layers
    compressed by 1.5 % and partly relaxed; its remaining overlaps (up to
    tens of um) push the bed apart within a few thousand steps, so it is
    not at rest.
    """
    rng = np.random.RandomState(seed)
    radius = np.clip(median * np.exp(sigma * rng.standard_normal(n)), rclip[0], rclip[1])
    lo = np.array([-0.5 * width, -0.5 * width, 0.0])
    wo, wn = hopper_walls(0.5 * width, angle_deg)
    if packing == "settled":
        pos = deposit_packing(radius, lo, (0.5 * width, 0.5 * width), wo, wn, seed=seed * 7919 + 11,
                              preload=1e-4)
    elif packing == "dense":
        radius, pos = layered_lattice(radius, lo, (0.5 * width, 0.5 * width), factor=0.985)
        relax_packing(radius, pos, lo, (0.5 * width, 0.5 * width), max_iter=60, tol=2e-3)
    else:
        radius = np.sort(radius)[::-1].copy()
        vol = float(np.sum(4.0 / 3.0 * np.pi * radius ** 3))
        height = vol / (fill * width * width)
        hi = np.array([0.5 * width, 0.5 * width, height])
        pos = rsa_positions(radius, lo, hi, seed=seed * 104729 + 3, max_attempts=20000)
    # input row order carries no spatial locality (the solver sorts by cell)
    perm = np.random.RandomState(seed + 1).permutation(n)
    radius, pos = radius[perm].copy(), pos[perm].copy()
    kN = E * median
    m = mech(epsS=restitution, mu=mu, epsN=restitution, cohesion=0.0, beta=beta, muT=muT, muR=muR,
             kN=kN)
    return Scene(**sphere_particles(pos, radius, density), wall_origin=wo, wall_normal=wn,
                 pp=m, pt=dict(m), pw=dict(m), dt=dt, gravity=tuple(gravity), gravity_enabled=1,
                 mesh_scale=mesh_scale, steps=steps)


class SceneBatch:
    """Independent scenes stepped by one solver (include/dem_hip.h,
    dem_initialize_batch).  The scenes share dt, gravity and the cell-size
    factor; each keeps its own particles, walls and particle-particle /
    particle-wall laws.  `combined` is the back-to-back Scene handed to the
    solver; `split` cuts a downloaded per-particle array back into scenes."""

    def __init__(self, scenes):
        scenes = list(scenes)
        if not scenes:
            raise ValueError("a batch needs at least one scene")
        s0 = scenes[0]
        for sc in scenes:
            if len(sc.body_mass):
                raise ValueError("batched scenes cannot carry bodies")
            if (sc.dt != s0.dt or tuple(sc.gravity) != tuple(s0.gravity)
                    or sc.gravity_enabled != s0.gravity_enabled or sc.mesh_scale != s0.mesh_scale
                    or sc.universal_gravity):
                raise ValueError("batched scenes must share dt, gravity and mesh_scale "
                                 "and cannot use universal gravitation")
        self.scenes = scenes
        self.particle_start = np.concatenate([[0], np.cumsum([sc.n for sc in scenes])]).astype(np.int64)
        self.wall_start = np.concatenate([[0], np.cumsum([len(sc.wall_origin) for sc in scenes])]
                                         ).astype(np.int32)
        cat = lambda name: np.concatenate([np.asarray(getattr(sc, name)) for sc in scenes])
        owner = np.concatenate([np.asarray(sc.comp_owner) + self.particle_start[k]
                                for k, sc in enumerate(scenes)]).astype(np.int32)
        self.combined = Scene(
            pid=cat("pid"), status=cat("status"), mass=cat("mass"), radius=cat("radius"),
            inertia=cat("inertia"), pos=cat("pos"), vel=cat("vel"), angvel=cat("angvel"),
            quat=cat("quat"), comp_id=cat("comp_id"), comp_owner=owner,
            comp_radius=cat("comp_radius"), comp_pos_body=cat("comp_pos_body"),
            wall_origin=cat("wall_origin").reshape(-1, 3), wall_normal=cat("wall_normal").reshape(-1, 3),
            pp=dict(s0.pp), pt=dict(s0.pt), pw=dict(s0.pw), dt=s0.dt, gravity=tuple(s0.gravity),
            gravity_enabled=s0.gravity_enabled, mesh_scale=s0.mesh_scale, steps=s0.steps)

    def __len__(self):
        return len(self.scenes)

    @property
    def n(self):
        return int(self.particle_start[-1])

    def split(self, a):
        return [a[self.particle_start[k]:self.particle_start[k + 1]] for k in range(len(self.scenes))]


def sweep_scenes(count=256, seed=4, nmin=500, nmax=20000, rmin=1.0e-3, rmax=1.5e-3, density=2500.0,
                 dt=1.0e-6, steps=2000, fill=0.2, mesh_scale=0.0, gravity=(0.0, 0.0, -9.81)):
    """A parameter sweep: `count` boxes, N ~ U[nmin, nmax] spheres of radius
    U[rmin, rmax], seeded RSA at volume fraction `fill` into a box twice as
    tall as wide, 6 walls; per scene mu ~ U[0.1, 0.9], restitution
    ~ U[0.2, 0.9], E log-uniform in [1e6, 1e8] (kN = E * mean radius),
    muT = muR = 0.1, beta = 0.5, no cohesion."""
    rng = np.random.RandomState(seed)
    rbar = 0.5 * (rmin + rmax)
    mean_r3 = (rmax ** 4 - rmin ** 4) / (4.0 * (rmax - rmin))
    out = []
    for k in range(count):
        n = int(rng.randint(nmin, nmax + 1))
        mu = float(rng.uniform(0.1, 0.9))
        e = float(rng.uniform(0.2, 0.9))
        E = float(10.0 ** rng.uniform(6.0, 8.0))
        radius = rng.uniform(rmin, rmax, n)
        vol = n * 4.0 / 3.0 * np.pi * mean_r3 / fill
        side = (vol / 2.0) ** (1.0 / 3.0)
        lo, hi = np.zeros(3), np.array([side, side, 2.0 * side])
        pos = rsa_positions(radius, lo, hi, seed=seed * 1_000_003 + k, max_attempts=20000)
        m = mech(epsS=e, mu=mu, epsN=e, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=E * rbar)
        wo, wn = box_walls(lo, hi)
        out.append(Scene(**sphere_particles(pos, radius, density), wall_origin=wo, wall_normal=wn,
                         pp=m, pt=dict(m), pw=dict(m), dt=dt, gravity=tuple(gravity), gravity_enabled=1,
                         mesh_scale=mesh_scale, steps=steps))
    return out


def settle_scenes(count=64, n=5000, seed=5, box=0.03, rmin=1.0e-3, rmax=1.5e-3, density=2500.0,
                  dt=1.0e-6, steps=2000, overlap=0.01, mesh_scale=0.0, gravity=(0.0, 0.0, -9.81)):
    """Settling beds for differentiation: `count` boxes of `box` x `box`
    footprint, n spheres of radius U[rmin, rmax] laid in layers (largest
    first) at spacing (1 - overlap) x diameter and partly relaxed, so the bed
    starts with small overlaps and settles under gravity against the floor
    (wall index 4 of box_walls); per scene mu ~ U[0.1, 0.9], restitution
    ~ U[0.2, 0.9], E log-uniform in [1e6, 1e8] (kN = E * mean radius),
    muT = muR = 0.1, beta = 0.5, no cohesion.  Row order is shuffled."""
    rng = np.random.RandomState(seed)
    out = []
    for k in range(count):
        mu = float(rng.uniform(0.1, 0.9))
        e = float(rng.uniform(0.2, 0.9))
        E = float(10.0 ** rng.uniform(6.0, 8.0))
        radius = rng.uniform(rmin, rmax, n)
        lo = np.zeros(3)
        radius, pos = layered_lattice(radius, lo, (box, box), factor=1.0 - overlap)
        relax_packing(radius, pos, lo, (box, box), max_iter=3, tol=overlap)
        perm = rng.permutation(n)
        radius, pos = radius[perm].copy(), pos[perm].copy()
        top = float(np.max(pos[:, 2] + radius))
        hi = np.array([box, box, top + 4.0 * rmax])
        m = mech(epsS=e, mu=mu, epsN=e, cohesion=0.0, beta=0.5, muT=0.1, muR=0.1, kN=E * float(np.mean(radius)))
        wo, wn = box_walls(lo, hi)
        out.append(Scene(**sphere_particles(pos, radius, density), wall_origin=wo, wall_normal=wn,
                         pp=m, pt=dict(m), pw=dict(m), dt=dt, gravity=tuple(gravity), gravity_enabled=1,
                         mesh_scale=mesh_scale, steps=steps))
    return out


# ---------------------------------------------------------------------------
# Readers for the reference's input formats.
# ---------------------------------------------------------------------------

def _strip_comment(line):
    return line.split("#", 1)[0].strip()


def read_world_par(path):
    """KEY = value pairs with '#' comments (reference simulation.cpp:158-307)."""
    out = {}
    with open(path, encoding="utf-8", errors="replace") as f:
        for raw in f:
            line = _strip_comment(raw)
            if not line or "=" not in line:
                continue
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def mechs_from_par(par):
    """MECH_ keys for all three blocks plus MECH_PP_/PT_/PW_ overrides
    (reference simulation.cpp:24-55)."""
    names = dict(epsS="epsS", mu="mu", epsN="epsN", cohesive="cohesion", beta="beta",
                 mu_T="muT", mu_R="muR", kN="kN")
    blocks = {"pp": mech(), "pt": mech(), "pw": mech()}
    for key, value in par.items():
        mobj = re.match(r"MECH_(PP_|PT_|PW_)?(\w+)$", key)
        if not mobj or mobj.group(2) not in names:
            continue
        field = names[mobj.group(2)]
        targets = ["pp", "pt", "pw"] if mobj.group(1) is None else [mobj.group(1)[:2].lower()]
        for t in targets:
            blocks[t][field] = float(value)
    return blocks


def read_particles_bt(path):
    """17 whitespace-separated fields per particle (reference particle.cpp:66-73)."""
    data = np.loadtxt(path, ndmin=2)
    pid = data[:, 0].astype(np.int32)
    status = data[:, 1].astype(np.int32)
    mass, radius = data[:, 2].copy(), data[:, 3].copy()
    inertia = 0.4 * mass * radius * radius
    return dict(pid=pid, status=status, mass=mass, radius=radius,
                inertia=np.repeat(inertia[:, None], 3, axis=1), pos=data[:, 4:7].copy(),
                vel=data[:, 7:10].copy(), quat=data[:, 10:14].copy(), angvel=data[:, 14:17].copy())


def _sidecar_records(path, what):
    """Non-blank lines of a sidecar with their 1-based line numbers."""
    if not os.path.isfile(path):
        raise RuntimeError(f"Cannot open {what}: {os.path.abspath(path)}")
    out = []
    with open(path, encoding="utf-8", errors="replace") as f:
        for number, line in enumerate(f, start=1):
            if line.strip(" \t\r\n"):
                out.append((number, line.split()))
    return out


def _int_token(tok):
    if not re.fullmatch(r"[+-]?\d+", tok):
        raise ValueError(tok)
    v = int(tok)
    if v < -2 ** 31 or v >= 2 ** 31:
        raise ValueError(tok)
    return v


def read_principal_inertia(path, pid):
    """PARTICLE::LoadPrincipalInertia (reference particle.cpp:108-163): one
    "owner_id Ixx Iyy Izz" record per owner (blank lines skipped), every
    owner exactly once, finite positive moments.  Returns [n][3] by owner row."""
    rows = {int(v): i for i, v in enumerate(pid)}
    inertia = np.zeros((len(pid), 3))
    seen = np.zeros(len(pid), bool)
    for number, tok in _sidecar_records(path, "particle principal-inertia sidecar"):
        try:
            if len(tok) != 4:
                raise ValueError
            owner = _int_token(tok[0])
            vals = [float(t) for t in tok[1:]]
        except ValueError:
            raise RuntimeError(f"Invalid principal-inertia record at {os.path.abspath(path)}:{number}; "
                               "expected owner_id Ixx Iyy Izz")
        if owner not in rows:
            raise RuntimeError(f"Unknown owner ID in principal-inertia sidecar: {owner}")
        r = rows[owner]
        if seen[r]:
            raise RuntimeError(f"Duplicate owner ID in principal-inertia sidecar: {owner}")
        if not all(np.isfinite(v) and v > 0.0 for v in vals):
            raise RuntimeError(f"Principal inertia must be finite and strictly positive for owner {owner}")
        inertia[r] = vals
        seen[r] = True
    if seen.sum() != len(pid):
        raise RuntimeError("Principal-inertia sidecar must contain exactly one record per owner")
    return inertia


def read_clump_components(path, pid):
    """PARTICLE::LoadClumpComponents (reference particle.cpp:165-278):
    "component_id owner_id radius body_x body_y body_z" per component, in
    file order; unique non-negative component ids, finite geometry with a
    positive radius, 1 to 8 components per owner.  Returns comp_id,
    comp_owner (owner row), comp_radius, comp_pos_body."""
    rows = {int(v): i for i, v in enumerate(pid)}
    recs = _sidecar_records(path, "clump-component sidecar")
    if not recs:
        raise RuntimeError("Clump-component sidecar must not be empty")
    cid, owner, radius, posb = [], [], [], []
    ids = set()
    counts = np.zeros(len(pid), int)
    for number, tok in recs:
        try:
            if len(tok) != 6:
                raise ValueError
            c, o = _int_token(tok[0]), _int_token(tok[1])
            r, x, y, z = (float(t) for t in tok[2:])
        except ValueError:
            raise RuntimeError(f"Invalid clump-component record at {os.path.abspath(path)}:{number}; "
                               "expected component_id owner_id radius body_x body_y body_z")
        if o not in rows:
            raise RuntimeError(f"Unknown owner ID in clump-component sidecar: {o}")
        if c in ids:
            raise RuntimeError(f"Duplicate component ID in clump-component sidecar: {c}")
        ids.add(c)
        if c < 0:
            raise RuntimeError("Clump component IDs must lie in [0, INT32_MAX]")
        if not (np.isfinite(r) and r > 0.0 and np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
            raise RuntimeError("Clump-component geometry must be finite with positive radius")
        cid.append(c); owner.append(rows[o]); radius.append(r); posb.append((x, y, z))
        counts[rows[o]] += 1
    for i, k in enumerate(counts):
        if k == 0:
            raise RuntimeError("Every particle owner must have at least one clump component; "
                               f"missing owner {int(pid[i])}")
        if k > 8:
            raise RuntimeError(f"Each particle owner may have at most 8 clump components; owner {int(pid[i])}")
    return dict(comp_id=np.array(cid, np.int32), comp_owner=np.array(owner, np.int32),
                comp_radius=np.array(radius, float), comp_pos_body=np.array(posb, float).reshape(-1, 3))


def read_walls_bt(path):
    tok = open(path).read().split()
    n = int(tok[0])
    vals = np.array(tok[1:1 + 6 * n], float).reshape(n, 6)
    return vals[:, :3].copy(), vals[:, 3:].copy()


def mesh_rows(pts, tri):
    """Per triangle (0-based vertex indices) the inner apex v1 = centroid -
    normal * mean edge (reference multibody.cpp:54-145).  Returns [t][15] =
    v1, v2, v3, v4, unit normal (v3 - v2) x (v4 - v2) in the body frame."""
    pts = np.asarray(pts, float)
    tri = np.asarray(tri, int)
    out = np.zeros((len(tri), 15))
    for k in range(len(tri)):
        v2, v3, v4 = pts[tri[k, 0]], pts[tri[k, 1]], pts[tri[k, 2]]
        a1, a2, a3 = v3 - v2, v4 - v2, v3 - v4
        center = (v2 + v3 + v4) / 3
        n1, n2, n3 = (float(np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])) for a in (a1, a2, a3))
        mean = (n1 + n2 + n3) / 3
        c = np.array([a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                      a1[0] * a2[1] - a1[1] * a2[0]])
        length = float(np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]))
        normal = c * (1.0 / length) if length > 1e-14 else np.zeros(3)
        v1 = center - normal * mean
        out[k] = np.concatenate([v1, v2, v3, v4, normal])
    return out


def read_mesh_bt(path):
    """Vertex count, triangle count, vertices, 1-based faces (reference
    multibody.cpp:54-145); see mesh_rows."""
    tok = open(path).read().split()
    nv, nt = int(tok[0]), int(tok[1])
    pts = np.array(tok[2:2 + 3 * nv], float).reshape(nv, 3)
    tri = np.array(tok[2 + 3 * nv:2 + 3 * nv + 3 * nt], int).reshape(nt, 3) - 1
    return mesh_rows(pts, tri)


def rectangle_mesh(origin, u, v, nu, nv):
    """origin + [0,1] u + [0,1] v split into nu x nv squares of two triangles,
    each facing u x v (the side a contact must come from).  Returns mesh rows."""
    origin, u, v = (np.asarray(a, float) for a in (origin, u, v))
    pts = np.array([origin + u * (i / nu) + v * (j / nv) for j in range(nv + 1) for i in range(nu + 1)])
    idx = lambda i, j: j * (nu + 1) + i
    tri = []
    for j in range(nv):
        for i in range(nu):
            tri.append((idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)))
            tri.append((idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)))
    return mesh_rows(pts, np.array(tri))


def mixer_scene(n=200_000, seed=6, side=0.22, rmin=1.0e-3, rmax=2.0e-3, density=2500.0, E=1.0e7, mu=0.5,
                restitution=0.5, dt=1.0e-6, steps=3000, muT=0.1, muR=0.1, beta=0.5, fill=0.3,
                gravity=(0.0, 0.0, -9.81), blade_clear=2.0e-3, plate_gap=3.0e-3, packing="rsa",
                snap=0.5, press=1.0e-3):
    """DEM-MBD coupled mixer: n spheres, radius U[rmin, rmax], seeded RSA in a
    square bin, and three triangle-mesh bodies:
      * the bin (fixed, state 1): floor 32 x 32 x 2 and four walls 16 x 16 x 2
        triangles (4,096), all facing inward;
      * a paddle (prescribed, state 2) turning at omega = (0, 0, 6.28) rad/s
        about the bin axis: a vertical blade, both faces 16 x 32 x 2 triangles
        (2,048);
      * a plate (dynamic, state 0, 1 kg, identity inertia) resting above the
        bed, 16 x 32 x 2 triangles (1,024) facing down.
    packing="rsa": seeded random sequential insertion at volume fraction
    `fill`; spheres closer than blade_clear to the blade are left out and the
    plate starts plate_gap above the highest sphere (no initial contacts).
    packing="dense": a settled-bed stand-in; each side of the blade plane
    y = 0 gets half the spheres as a jittered lattice relaxed to near
    contact (as the dense hopper), then every sphere within snap * r of the
    floor, a bin wall or the blade plane is moved onto it with an overlap of
    press * r, and the plate rests on the highest spheres under it, those
    within snap * r of it pressed against it the same way.
    kN = E * mean radius, epsS = epsN = restitution, no cohesion."""
    rng = np.random.RandomState(seed)
    h = 1.5 * side
    # bin: floor at z = 0, walls to z = h, origin at the floor centre
    a = 0.5 * side
    bin_rows = np.concatenate([
        rectangle_mesh((-a, -a, 0.0), (side, 0, 0), (0, side, 0), 32, 32),          # floor, +z
        rectangle_mesh((-a, -a, 0.0), (0, side, 0), (0, 0, h), 16, 16),             # x = -a, +x
        rectangle_mesh((a, -a, 0.0), (0, 0, h), (0, side, 0), 16, 16),              # x = +a, -x
        rectangle_mesh((-a, -a, 0.0), (0, 0, h), (side, 0, 0), 16, 16),             # y = -a, +y
        rectangle_mesh((-a, a, 0.0), (side, 0, 0), (0, 0, h), 16, 16)])             # y = +a, -y
    # paddle blade in the plane y = 0 of its body frame, centre at (0, 0, zc)
    bw, bh, zc = 0.8 * side, 0.25 * side, 0.15 * side
    paddle_rows = np.concatenate([
        rectangle_mesh((-0.5 * bw, 0.0, -0.5 * bh), (bw, 0, 0), (0, 0, bh), 16, 32),   # faces -y
        rectangle_mesh((-0.5 * bw, 0.0, -0.5 * bh), (0, 0, bh), (bw, 0, 0), 32, 16)])  # faces +y
    pw, pl = 0.4 * side, 0.8 * side
    if packing == "dense":
        radius = rng.uniform(rmin, rmax, n)
        halves = []
        for k, (ylo, yhi) in enumerate(((-a, 0.0), (0.0, a))):
            lo = np.array([-a, ylo, 0.0])
            r, p = layered_lattice(radius[k::2], lo, (a, yhi), factor=0.985)
            relax_packing(r, p, lo, (a, yhi), max_iter=60, tol=2e-3)
            for axis, c, inward in ((0, -a, 1.0), (0, a, -1.0), (1, ylo, 1.0), (1, yhi, -1.0), (2, 0.0, 1.0)):
                near = inward * (p[:, axis] - c) - r < snap * r
                p[near, axis] = c + inward * r[near] * (1.0 - press)
            halves.append((r, p))
        radius = np.concatenate([halves[0][0], halves[1][0]])
        pos = np.concatenate([halves[0][1], halves[1][1]])
        if np.max(pos[:, 2] + radius) > h:
            raise ValueError("the bin is too small for the bed")
        under = (np.abs(pos[:, 0]) < 0.5 * pl) & (np.abs(pos[:, 1]) < 0.5 * pw)
        plate_z = float(np.max(pos[under, 2] + radius[under]))
        lift = under & (plate_z - pos[:, 2] - radius < snap * radius)
        pos[lift, 2] = plate_z - radius[lift] * (1.0 - press)
        perm = np.random.RandomState(seed + 1).permutation(n)
        radius, pos = radius[perm].copy(), pos[perm].copy()
    else:
        # spheres: RSA in the bin, leaving out those that cut the blade slab
        radius = np.sort(rng.uniform(rmin, rmax, int(n * 1.08) + 64))[::-1].copy()  # largest first
        vol = float(np.sum(4.0 / 3.0 * np.pi * radius ** 3))
        top = vol / (fill * side * side)
        if top > h:
            raise ValueError("the bin is too small for the bed")
        lo = np.array([-a, -a, 0.0])
        hi = np.array([a, a, top])
        pos = rsa_positions(radius, lo, hi, seed=seed * 7919 + 5, max_attempts=20000)
        perm = np.random.RandomState(seed + 1).permutation(len(radius))
        radius, pos = radius[perm], pos[perm]
        clear = blade_clear
        cut = (np.abs(pos[:, 1]) < radius + clear) & (np.abs(pos[:, 0]) < 0.5 * bw + radius + clear) & \
              (np.abs(pos[:, 2] - zc) < 0.5 * bh + radius + clear)
        keep = np.nonzero(~cut)[0][:n]
        if len(keep) < n:
            raise RuntimeError("mixer packing placed too few spheres")
        radius, pos = radius[keep].copy(), pos[keep].copy()
        plate_z = float(np.max(pos[:, 2] + radius)) + plate_gap
    # plate on or above the bed, facing down, body centre at its middle
    plate_rows = rectangle_mesh((-0.5 * pl, -0.5 * pw, 0.0), (0, pw, 0), (pl, 0, 0), 16, 32)
    kN = E * 0.5 * (rmin + rmax)
    m = mech(epsS=restitution, mu=mu, epsN=restitution, cohesion=0.0, beta=beta, muT=muT, muR=muR, kN=kN)
    ident = [1.0, 0.0, 0.0, 0.0]
    pose = np.array([[0.0, 0.0, 0.0, *ident, 0, 0, 0, 0, 0, 0],
                     [0.0, 0.0, zc, *ident, 0, 0, 0, 0.0, 0.0, 6.28],
                     [0.0, 0.0, plate_z, *ident, 0, 0, 0, 0, 0, 0]])
    eye = np.eye(3)
    return Scene(**sphere_particles(pos, radius, density), pp=m, pt=dict(m), pw=dict(m), dt=dt,
                 gravity=tuple(gravity), gravity_enabled=1, steps=steps,
                 body_mass=np.array([1000.0, 10.0, 1.0]), body_pose=pose,
                 body_tri_count=np.array([len(bin_rows), len(paddle_rows), len(plate_rows)], np.int32),
                 tri_local=np.concatenate([bin_rows, paddle_rows, plate_rows]),
                 body_state=np.array([1, 2, 0], np.int32), body_inertia=np.stack([eye, eye, eye]))


def read_bodies_csv(path, resolve):
    """Count line, header, then 24 or 25 columns per body; the optional 25th
    column is the motion state 0/1/2 (reference multibody.cpp:295-435)."""
    lines = [ln.rstrip("\r\n") for ln in open(path, encoding="utf-8", errors="replace")]
    count = int(lines[0].split(",")[0])
    rows = [ln.split(",") for ln in lines[2:2 + count]]
    mass, pose, state, inertia, tris, counts = [], [], [], [], [], []
    for r in rows:
        vals = [float(x) for x in r[1:24]]
        st = int(r[24]) if len(r) >= 25 and r[24].strip() != "" else 0
        vel = vals[4:7] if st != 1 else [0.0, 0.0, 0.0]
        omg = vals[7:10] if st != 1 else [0.0, 0.0, 0.0]
        mass.append(vals[0])
        pose.append(vals[1:4] + vals[10:14] + vel + omg)
        inertia.append(np.array(vals[14:23]).reshape(3, 3))
        state.append(st)
        mesh = read_mesh_bt(resolve(r[0]))
        tris.append(mesh)
        counts.append(len(mesh))
    return dict(body_mass=np.array(mass), body_pose=np.array(pose),
                body_state=np.array(state, np.int32), body_inertia=np.array(inertia),
                tri_local=np.concatenate(tris) if tris else np.zeros((0, 15)),
                body_tri_count=np.array(counts, np.int32))


def write_case(sc, case_dir, steps, output_interval=None, outputs=True):
    """Writes a pure-DEM Scene (single-sphere owners, walls, no bodies) as a
    reference case directory: InputFile/Particles/Particles.bt (the 17 fields
    of particle.cpp:66-73, written with the ph-file writer, whose columns are
    those fields), InputFile/Environment/Wall.bt and
    InputFile/SettingData/world.par (simulation.cpp:158-307) with the
    scene's MECH block, gravity, step size and `steps` steps; frame files
    every `output_interval` steps when `outputs` is set, none otherwise."""
    from . import output

    if len(sc.body_mass) or len(sc.comp_id) != sc.n or np.any(sc.comp_pos_body != 0.0):
        raise ValueError("write_case takes pure-DEM scenes of single spheres")
    inp = os.path.join(case_dir, "InputFile")
    output.write_particle_state(os.path.join(inp, "Particles", "Particles.bt"), sc.pid, sc.status, sc.mass,
                                sc.radius, sc.pos, sc.vel, sc.quat, sc.angvel)
    os.makedirs(os.path.join(inp, "Environment"), exist_ok=True)
    with open(os.path.join(inp, "Environment", "Wall.bt"), "w") as f:
        f.write(f"{len(sc.wall_origin)}\n")
        for o, nrm in zip(np.reshape(sc.wall_origin, (-1, 3)), np.reshape(sc.wall_normal, (-1, 3))):
            f.write(" ".join("%.17g" % v for v in (*o, *nrm)) + "\n")
    m = sc.pp
    g = "%.17g"
    par = {"ODE_Option": 1, "ODE_StepSize": g % sc.dt, "ODE_EndStep": int(steps), "ODE_StartStep": 0,
           "ODE_OutputInterval": int(output_interval or steps or 1),
           "MECH_mu": g % m["mu"], "MECH_mu_T": g % m["muT"], "MECH_mu_R": g % m["muR"],
           "MECH_epsN": g % m["epsN"], "MECH_epsS": g % m["epsS"], "MECH_cohesive": g % m["cohesion"],
           "MECH_beta": g % m["beta"], "MECH_kN": g % m["kN"],
           "CONTROL_GRAVITY_FLAG": int(sc.gravity_enabled), "CONTROL_GRAVITY_X": g % sc.gravity[0],
           "CONTROL_GRAVITY_Y": g % sc.gravity[1], "CONTROL_GRAVITY_Z": g % sc.gravity[2],
           "CONTROL_WALL_FLAG": 1 if len(sc.wall_origin) else 0,
           "CONTROL_WALL_FILE": "Data\\DATA\\InputFile\\Environment\\Wall.bt", "CONTROL_MULTIBODY_FLAG": 0,
           "CONTROL_UNIVERSAL_GRAVITATION_FLAG": int(sc.universal_gravity),
           "CONTROL_PARTICLES_FILE": "Data\\DATA\\InputFile\\Particles\\Particles.bt"}
    for key in ("PARTICLE_STATE", "PARTICLE_FORCE", "BODY_STATE", "BODY_FORCE"):
        par[f"CONTROL_OUTPUT_{key}"] = 1 if outputs else 0
    par["CONTROL_OUTPUT_SUMMARY"] = 0
    os.makedirs(os.path.join(inp, "SettingData"), exist_ok=True)
    with open(os.path.join(inp, "SettingData", "world.par"), "w") as f:
        f.write("".join(f"{k} = {v}\n" for k, v in par.items()))
    return case_dir


def read_case(case_dir):
    """Loads a reference case directory (InputFile/...) into a Scene."""
    par = read_world_par(os.path.join(case_dir, "InputFile", "SettingData", "world.par"))

    def resolve(p):
        p = p.strip().replace("\\", "/")
        p = re.sub(r"^Data/DATA/", "", p)
        return os.path.join(case_dir, p)

    parts = read_particles_bt(resolve(par["CONTROL_PARTICLES_FILE"]))
    kw = dict(parts)
    # sidecars (reference simulation.cpp:109-110): explicit principal inertia
    # and clump components; absent or empty keys keep the solid-sphere defaults
    if par.get("CONTROL_PARTICLE_INERTIA_FILE", "").strip():
        kw["inertia"] = read_principal_inertia(resolve(par["CONTROL_PARTICLE_INERTIA_FILE"]), kw["pid"])
    if par.get("CONTROL_CLUMP_COMPONENTS_FILE", "").strip():
        kw.update(read_clump_components(resolve(par["CONTROL_CLUMP_COMPONENTS_FILE"]), kw["pid"]))
    elif np.any(kw["pid"] < 0):
        raise RuntimeError("Default component ID requires a non-negative particle ID")
    if int(par.get("CONTROL_WALL_FLAG", "0")):
        kw["wall_origin"], kw["wall_normal"] = read_walls_bt(resolve(par["CONTROL_WALL_FILE"]))
    if int(par.get("CONTROL_MULTIBODY_FLAG", "0")):
        kw.update(read_bodies_csv(resolve(par["CONTROL_MULTIBODY_FILE"]), resolve))
    blocks = mechs_from_par(par)
    g = [float(par.get(f"CONTROL_GRAVITY_{a}", "0")) for a in "XYZ"]
    return Scene(**kw, pp=blocks["pp"], pt=blocks["pt"], pw=blocks["pw"],
                 dt=float(par["ODE_StepSize"]), gravity=tuple(g),
                 gravity_enabled=1 if int(par.get("CONTROL_GRAVITY_FLAG", "0")) else 0,
                 universal_gravity=int(par.get("CONTROL_UNIVERSAL_GRAVITATION_FLAG", "0")),
                 keep_body_particle=int(par.get("CONTROL_OUTPUT_PARTICLE_FORCE", "1")),
                 keep_body_detail=int(par.get("CONTROL_OUTPUT_BODY_FORCE", "1")),
                 steps=int(par["ODE_EndStep"]) - int(par.get("ODE_StartStep", "0")),
                 output_interval=int(par.get("ODE_OutputInterval", "1")))
