"""ctypes declarations of include/dem_hip.h and a thin solver wrapper.

The wrapper binds one shared library whose entry points follow the layout of
include/dem_hip.h under a name prefix ("dem_" for libdemhip.so).  All arrays
crossing the boundary are C-contiguous numpy arrays owned by Python.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

DEM_PARTICLE_HISTORY_SLOTS = 64
DEM_WALL_HISTORY_SLOTS = 16
DEM_BODY_HISTORY_SLOTS = 64
DEM_BODY_REGIME_COUNTERS = 7
DEM_BODY_MICROSTATE_VALUES = 22

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)


class ParticlesView(C.Structure):
    _fields_ = [("count", C.c_int32), ("id", _ip), ("status", _ip), ("mass", _dp),
                ("radius", _dp), ("inertia", _dp), ("pos", _dp), ("vel", _dp),
                ("ang_vel", _dp), ("quat", _dp)]


class ComponentsView(C.Structure):
    _fields_ = [("count", C.c_int32), ("id", _ip), ("owner", _ip), ("radius", _dp),
                ("pos_body", _dp)]


class WallsView(C.Structure):
    _fields_ = [("count", C.c_int32), ("origin", _dp), ("normal", _dp)]


class BodiesView(C.Structure):
    _fields_ = [("count", C.c_int32), ("mass", _dp), ("pose13", _dp), ("tri_count", _ip),
                ("tri_local", _dp)]


class Mech(C.Structure):
    _fields_ = [(name, C.c_double) for name in
                ("epsS", "mu", "epsN", "cohesion", "beta", "muT", "muR", "kN")]


class Control(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("gravity_enabled", C.c_int32),
                ("universal_gravity", C.c_int32), ("keep_body_particle", C.c_int32),
                ("keep_body_detail", C.c_int32), ("mesh_scale", C.c_double)]


class BatchView(C.Structure):
    _fields_ = [("scene_count", C.c_int32), ("particle_start", C.POINTER(C.c_int64)),
                ("wall_start", _ip), ("pp", C.POINTER(Mech)), ("pw", C.POINTER(Mech))]


class StepStats(C.Structure):
    _fields_ = [("particle_contacts", C.c_int32), ("body_contacts", C.c_int32),
                ("body_max_depth", C.c_double), ("body_contact_fz", C.c_double),
                ("particle_history_overflows", C.c_int32), ("body_history_overflows", C.c_int32),
                ("particle_history_high_water", C.c_int32), ("body_history_high_water", C.c_int32),
                ("triangle_grid_overflows", C.c_int32), ("reserved", C.c_int32),
                ("body_regime_interval_native_steps", C.c_uint64)]


class DemError(RuntimeError):
    """Raised with the library's own message when an entry point fails."""


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _mech(m):
    return Mech(*[float(m[k]) for k in ("epsS", "mu", "epsN", "cohesion", "beta", "muT", "muR", "kN")])


class Library:
    """One loaded shared library exposing the dem_hip.h entry points."""

    def __init__(self, path: str, prefix: str = "dem_"):
        if not os.path.exists(path):
            raise DemError(f"shared library not found: {path}")
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        p = prefix
        L = self.lib
        self._fn = {}

        def bind(name, argtypes, restype=C.c_int):
            f = getattr(L, p + name)
            f.argtypes = argtypes
            f.restype = restype
            self._fn[name] = f

        vp = C.c_void_p
        bind("create", [C.POINTER(vp)])
        bind("destroy", [vp], None)
        bind("initialize", [vp, C.POINTER(ParticlesView), C.POINTER(ComponentsView),
                            C.POINTER(WallsView), C.POINTER(BodiesView), C.POINTER(Mech),
                            C.POINTER(Mech), C.POINTER(Mech), C.POINTER(Control)])
        if hasattr(L, p + "initialize_batch"):
            bind("initialize_batch", [vp, C.POINTER(ParticlesView), C.POINTER(ComponentsView),
                                      C.POINTER(WallsView), C.POINTER(BatchView), C.POINTER(Control)])
        bind("compute_forces", [vp, _dp, _dp, C.c_int])
        if hasattr(L, p + "compute_forces_noalloc"):
            bind("compute_forces_noalloc", [vp, _dp, _dp, C.c_int])
        bind("init_half_step", [vp, C.c_double])
        bind("advance_particles", [vp, C.c_double, _dp])
        bind("finish_particle_step", [vp, C.c_double, C.c_double])
        bind("download_particle_state", [vp, _dp, _dp, _dp, _dp])
        bind("download_force_output", [vp, _dp, _dp, _dp, _dp, _dp, _dp])
        bind("get_step_stats", [vp, C.POINTER(StepStats), _u64p, _u64p, _dp, _dp])
        bind("particle_grid_entry_count", [vp, C.POINTER(C.c_size_t)])
        bind("triangle_grid_entry_count", [vp, C.POINTER(C.c_size_t)])
        bind("debug_grid", [vp, _u64p, _ip])
        bind("debug_history", [vp, _u64p, _dp, _u64p, _dp, _u64p, _dp])
        bind("debug_triangle_grid", [vp, _u64p, _ip])
        bind("last_error", [], C.c_char_p)
        for opt in ("load_contact_history", "save_contact_history"):
            if hasattr(L, p + opt):
                bind(opt, [vp, C.c_char_p])
        if hasattr(L, p + "has_restored_leapfrog_state"):
            bind("has_restored_leapfrog_state", [vp, C.POINTER(C.c_int)])
        if hasattr(L, p + "device_memory_bytes"):
            bind("device_memory_bytes", [vp, C.POINTER(C.c_size_t)])
        if hasattr(L, p + "backend_available"):
            bind("backend_available", [])
        if hasattr(L, p + "device_description"):
            bind("device_description", [], C.c_char_p)
        if hasattr(L, p + "synchronize"):
            bind("synchronize", [vp])
        if hasattr(L, p + "run_steps"):
            bind("run_steps", [vp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)])
        if hasattr(L, p + "diff_forward"):
            bind("diff_forward", [vp, C.c_int, C.c_int, C.c_int, _dp, _dp])
            bind("diff_backward", [vp, _dp, _dp, _dp])
        if hasattr(L, p + "debug_last_timings"):
            bind("debug_last_timings", [vp, _dp, C.c_int, C.POINTER(C.c_int)])
        if hasattr(L, p + "set_threads"):
            bind("set_threads", [vp, C.c_int])
        if hasattr(L, p + "debug_run_counters"):
            bind("debug_run_counters", [vp, C.POINTER(C.c_longlong), C.c_int, C.POINTER(C.c_int)])
        if hasattr(L, p + "debug_enable_timings"):
            bind("debug_enable_timings", [vp, C.c_int])

    def call(self, name, *args):
        rc = self._fn[name](*args)
        if rc != 0:
            msg = self._fn["last_error"]()
            raise DemError(msg.decode() if msg else f"{self.prefix}{name} failed ({rc})")
        return rc

    def has(self, name):
        return name in self._fn

    def solver(self) -> "Solver":
        return Solver(self)


class Solver:
    """Python mirror of the reference's CudaDemSolver over a dem_hip.h library."""

    def __init__(self, lib: Library):
        self.lib = lib
        h = C.c_void_p()
        lib.call("create", C.byref(h))
        self.h = h
        self.n = 0
        self.nc = 0
        self.nb = 0
        self._keep = []
        # compute_forces goes through dem_compute_forces_noalloc when set
        self.noalloc = False

    def close(self):
        if self.h:
            self.lib._fn["destroy"](self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _views(self, s):
        n = int(len(s.pid))
        a = dict(
            pid=_i32(s.pid), status=_i32(s.status), mass=_f64(s.mass), radius=_f64(s.radius),
            inertia=_f64(s.inertia, (n, 3)), pos=_f64(s.pos, (n, 3)), vel=_f64(s.vel, (n, 3)),
            angvel=_f64(s.angvel, (n, 3)), quat=_f64(s.quat, (n, 4)),
            cid=_i32(s.comp_id), cowner=_i32(s.comp_owner), cr=_f64(s.comp_radius),
            cpb=_f64(s.comp_pos_body, (-1, 3)),
            wo=_f64(s.wall_origin, (-1, 3)), wn=_f64(s.wall_normal, (-1, 3)),
            bm=_f64(s.body_mass), bp=_f64(s.body_pose, (-1, 13)), btc=_i32(s.body_tri_count),
            tl=_f64(s.tri_local, (-1, 15)))
        self._keep.append(a)
        pv = ParticlesView(n, _ptr(a["pid"], C.c_int32), _ptr(a["status"], C.c_int32),
                           _ptr(a["mass"], C.c_double), _ptr(a["radius"], C.c_double),
                           _ptr(a["inertia"], C.c_double), _ptr(a["pos"], C.c_double),
                           _ptr(a["vel"], C.c_double), _ptr(a["angvel"], C.c_double),
                           _ptr(a["quat"], C.c_double))
        cv = ComponentsView(len(a["cid"]), _ptr(a["cid"], C.c_int32), _ptr(a["cowner"], C.c_int32),
                            _ptr(a["cr"], C.c_double), _ptr(a["cpb"], C.c_double))
        wv = WallsView(len(a["wo"]), _ptr(a["wo"], C.c_double), _ptr(a["wn"], C.c_double))
        bv = BodiesView(len(a["bm"]), _ptr(a["bm"], C.c_double), _ptr(a["bp"], C.c_double),
                        _ptr(a["btc"], C.c_int32), _ptr(a["tl"], C.c_double))
        ctl = Control(float(s.dt), (C.c_double * 3)(*[float(g) for g in s.gravity]),
                      int(s.gravity_enabled), int(s.universal_gravity),
                      int(s.keep_body_particle), int(s.keep_body_detail), float(s.mesh_scale))
        return n, a, pv, cv, wv, bv, ctl

    def _after_init(self, s, n, a):
        self.n = n
        self.nc = len(a["cid"])
        self.nb = len(a["bm"])
        self.keep_body_particle = bool(s.keep_body_particle)
        self.keep_body_detail = bool(s.keep_body_detail)

    def initialize(self, scene):
        s = scene
        self._keep = []
        n, a, pv, cv, wv, bv, ctl = self._views(s)
        mechs = [_mech(m) for m in (s.pp, s.pt, s.pw)]
        self.lib.call("initialize", self.h, C.byref(pv), C.byref(cv), C.byref(wv), C.byref(bv),
                      C.byref(mechs[0]), C.byref(mechs[1]), C.byref(mechs[2]), C.byref(ctl))
        self._after_init(s, n, a)

    def initialize_batch(self, batch):
        """Initialises the solver with a SceneBatch (scenes back to back)."""
        self._keep = []
        s = batch.combined
        n, a, pv, cv, wv, bv, ctl = self._views(s)
        k = len(batch.scenes)
        rows = np.ascontiguousarray(batch.particle_start, dtype=np.int64)
        walls = _i32(batch.wall_start)
        pp = (Mech * k)(*[_mech(sc.pp) for sc in batch.scenes])
        pw = (Mech * k)(*[_mech(sc.pw) for sc in batch.scenes])
        self._keep.append((rows, walls, pp, pw))
        view = BatchView(k, _ptr(rows, C.c_int64), _ptr(walls, C.c_int32), pp, pw)
        self.lib.call("initialize_batch", self.h, C.byref(pv), C.byref(cv), C.byref(wv), C.byref(view),
                      C.byref(ctl))
        self._after_init(s, n, a)

    def compute_forces(self, body_pose=None, loads=False, collect=False):
        pose = None if body_pose is None else _f64(body_pose, (self.nb, 13))
        out = np.zeros((self.nb, 6)) if loads else None
        self.lib.call("compute_forces_noalloc" if self.noalloc else "compute_forces", self.h,
                      _ptr(pose, C.c_double), _ptr(out, C.c_double), 1 if collect else 0)
        return out

    def init_half_step(self, half_dt):
        self.lib.call("init_half_step", self.h, float(half_dt))

    def advance(self, dt):
        v = C.c_double(-1.0)
        self.lib.call("advance_particles", self.h, float(dt), C.byref(v))
        return v.value

    def finish(self, dt, half_dt=None):
        self.lib.call("finish_particle_step", self.h, float(dt),
                      float(0.5 * dt if half_dt is None else half_dt))

    def state(self):
        n = self.n
        pos, vel, ang, quat = (np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)),
                               np.zeros((n, 4)))
        self.lib.call("download_particle_state", self.h, _ptr(pos, C.c_double),
                      _ptr(vel, C.c_double), _ptr(ang, C.c_double), _ptr(quat, C.c_double))
        return dict(pos=pos, vel=vel, angvel=ang, quat=quat)

    def forces(self, body_particle=None, body_detail=None):
        n, nb = self.n, self.nb
        bp_on = self.keep_body_particle if body_particle is None else body_particle
        bd_on = self.keep_body_detail if body_detail is None else body_detail
        out = {k: np.zeros((n, 6)) for k in ("contact", "field", "wall", "total")}
        bp = np.zeros((nb, n, 6)) if (bp_on and nb) else None
        bd = np.zeros((nb, n, 19)) if (bd_on and nb) else None
        self.lib.call("download_force_output", self.h, _ptr(out["contact"], C.c_double),
                      _ptr(out["field"], C.c_double), _ptr(out["wall"], C.c_double),
                      _ptr(out["total"], C.c_double), _ptr(bp, C.c_double),
                      _ptr(bd, C.c_double))
        out["body_particle"] = bp
        out["body_detail"] = bd
        return out

    def stats(self):
        st = StepStats()
        nb = self.nb
        ri = np.zeros((nb, 7), np.uint64)
        rv = np.zeros((nb, 7), np.uint64)
        mi = np.zeros((nb, 22))
        mv = np.zeros((nb, 22))
        self.lib.call("get_step_stats", self.h, C.byref(st), _ptr(ri, C.c_uint64),
                      _ptr(rv, C.c_uint64), _ptr(mi, C.c_double), _ptr(mv, C.c_double))
        d = {name: getattr(st, name) for name, _ in StepStats._fields_}
        d.update(regime_instant=ri, regime_interval=rv, micro_instant=mi, micro_interval=mv)
        return d

    def grid_entry_counts(self):
        a, b = C.c_size_t(), C.c_size_t()
        self.lib.call("particle_grid_entry_count", self.h, C.byref(a))
        self.lib.call("triangle_grid_entry_count", self.h, C.byref(b))
        return a.value, b.value

    def debug_grid(self):
        keys = np.zeros(self.nc, np.uint64)
        ids = np.zeros(self.nc, np.int32)
        self.lib.call("debug_grid", self.h, _ptr(keys, C.c_uint64), _ptr(ids, C.c_int32))
        return keys, ids

    def set_threads(self, threads):
        """Owner-loop threads of a library that offers them (the oracle)."""
        if self.lib.has("set_threads"):
            self.lib.call("set_threads", self.h, int(threads))

    def run_counters(self):
        """dem_debug_run_counters as a dict (None for a library without it)."""
        if not self.lib.has("debug_run_counters"):
            return None
        buf = (C.c_longlong * 9)()
        cnt = C.c_int(0)
        self.lib.call("debug_run_counters", self.h, buf, 9, C.byref(cnt))
        names = ("evaluations", "sorts", "rebuilds", "graph_steps", "eager_steps", "last_sorted",
                 "scan_lds_passes", "scan_hbm_passes", "skin_pm")
        return {names[i]: int(buf[i]) for i in range(cnt.value)}

    def debug_history(self):
        nc = self.nc
        nt = 64 if self.nb else 0  # no body table without bodies (NULL: not read)
        out = dict(pp_keys=np.zeros((nc, 64), np.uint64), pp_values=np.zeros((nc, 64, 9)),
                   pw_keys=np.zeros((nc, 16), np.uint64), pw_values=np.zeros((nc, 16, 9)),
                   pt_keys=np.zeros((nc, nt), np.uint64), pt_values=np.zeros((nc, nt, 9)))
        if not nt:
            self.lib.call("debug_history", self.h, _ptr(out["pp_keys"], C.c_uint64),
                          _ptr(out["pp_values"], C.c_double), _ptr(out["pw_keys"], C.c_uint64),
                          _ptr(out["pw_values"], C.c_double), None, None)
            return out
        self.lib.call("debug_history", self.h, _ptr(out["pp_keys"], C.c_uint64),
                      _ptr(out["pp_values"], C.c_double), _ptr(out["pw_keys"], C.c_uint64),
                      _ptr(out["pw_values"], C.c_double), _ptr(out["pt_keys"], C.c_uint64),
                      _ptr(out["pt_values"], C.c_double))
        return out

    def debug_triangle_grid(self):
        _, m = self.grid_entry_counts()
        keys = np.zeros(m, np.uint64)
        tris = np.zeros(m, np.int32)
        self.lib.call("debug_triangle_grid", self.h, _ptr(keys, C.c_uint64), _ptr(tris, C.c_int32))
        return keys, tris

    def save_history(self, path):
        self.lib.call("save_contact_history", self.h, str(path).encode())

    def load_history(self, path):
        self.lib.call("load_contact_history", self.h, str(path).encode())

    def restored_leapfrog(self):
        v = C.c_int(0)
        self.lib.call("has_restored_leapfrog_state", self.h, C.byref(v))
        return bool(v.value)

    def device_memory_bytes(self):
        v = C.c_size_t(0)
        self.lib.call("device_memory_bytes", self.h, C.byref(v))
        return v.value

    def synchronize(self):
        if self.lib.has("synchronize"):
            self.lib.call("synchronize", self.h)

    def run_steps(self, steps, dt, graph=True, autotune=False):
        """`steps` x (advance, compute_forces without diagnostics, finish) on the
        device (HIP graphs unless graph=False).  Libraries without the entry
        point run the three calls per step."""
        if not self.lib.has("run_steps"):
            for _ in range(steps):
                self.advance(dt)
                self.compute_forces(collect=False)
                self.finish(dt)
            return steps
        done = C.c_int(0)
        flags = (0 if graph else 1) | (2 if autotune else 0)
        self.lib.call("run_steps", self.h, int(steps), float(dt), flags, C.byref(done))
        return done.value

    def diff_forward(self, steps, checkpoint_interval, plate_wall, scenes=1):
        """Differentiable run; returns per-scene (bed height, plate load)."""
        h = np.zeros(scenes)
        f = np.zeros(scenes)
        self.lib.call("diff_forward", self.h, int(steps), int(checkpoint_interval), int(plate_wall),
                      _ptr(h, C.c_double), _ptr(f, C.c_double))
        return h, f

    def diff_backward(self, d_height, d_load):
        """Per-scene d(objective)/d(MECH): [scenes][8] = pp (mu, epsN, epsS, kN),
        pw (mu, epsN, epsS, kN)."""
        dh = _f64(d_height)
        dl = _f64(d_load)
        out = np.zeros((len(dh), 8))
        self.lib.call("diff_backward", self.h, _ptr(dh, C.c_double), _ptr(dl, C.c_double),
                      _ptr(out, C.c_double))
        return out

    def enable_timings(self, on=True):
        self.lib.call("debug_enable_timings", self.h, 1 if on else 0)

    def last_timings(self):
        buf = np.zeros(16)
        cnt = C.c_int(0)
        self.lib.call("debug_last_timings", self.h, _ptr(buf, C.c_double), 16, C.byref(cnt))
        return buf[:cnt.value].copy()
