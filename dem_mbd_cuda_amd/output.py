"""The reference host's output files (AsyncOutputWriter, async_output.cpp;
writers in particle.cpp:320-610, multibody.cpp:173-228 and the constitutive-
branch / microstate CSVs of cuda_simulation_integrator.cpp:77-155), written
under <case>/OutputFile/ exactly as the reference names and formats them:

  OutputFile/state_particles/ph.%08d.bt     particle state
  OutputFile/force_particles/F.%08d.bt      particle forces (+ per-body FI_pt)
  OutputFile/force_bodys/FB.%05d.bt         per (body, particle) contact detail
  OutputFile/state_bodys/At.%08d.bt         body state, load and impulse
  OutputFile/native_body_contact_branches.csv
  OutputFile/native_body_contact_microstate.csv

Frame index = step + 1 (the initial frame, step -1, is index 0).  The frame
files are formatted by libdemio.so (printf conversions identical to the
reference's); a worker thread writes them while the solver keeps stepping,
as the reference's writer thread does.
"""
from __future__ import annotations

import ctypes as C
import os
import queue
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_io = None

MICROSTATE_NAMES = (
    "normal_load_sum_n", "normal_load_weighted_particle_radius_n_m", "normal_load_weighted_arm_radius_n_m",
    "twist_history_world_x_nm", "twist_history_world_y_nm", "twist_history_world_z_nm",
    "twist_history_norm_sum_nm", "twist_limit_sum_nm", "tangent_history_norm_sum_n", "tangent_limit_sum_n",
    "twist_utilization_area_sum", "tangent_utilization_area_sum", "contact_area_scale_sum",
    "normal_load_world_x_n", "normal_load_world_y_n", "normal_load_world_z_n",
    "tangent_history_world_x_n", "tangent_history_world_y_n", "tangent_history_world_z_n",
    "tangent_history_arm_torque_world_x_nm", "tangent_history_arm_torque_world_y_nm",
    "tangent_history_arm_torque_world_z_nm")

REGIME_HEADER = ("schema_version,end_step,end_time_s,body_id,window_native_steps,"
                 "instant_active,instant_primary_stick,instant_primary_slide,"
                 "instant_primary_twist,instant_slide_yield,instant_twist_yield,"
                 "instant_slide_and_twist,window_active,window_primary_stick,"
                 "window_primary_slide,window_primary_twist,window_slide_yield,"
                 "window_twist_yield,window_slide_and_twist\n")
MICRO_HEADER = ("schema_version,end_step,end_time_s,body_id,window_native_steps"
                + "".join(",instant_" + n for n in MICROSTATE_NAMES)
                + "".join(",window_" + n for n in MICROSTATE_NAMES) + "\n")


def io_lib():
    global _io
    if _io is None:
        path = os.path.join(_HERE, "libdemio.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run dem_mbd_cuda_amd.build.build_all()")
        L = C.CDLL(path)
        dp, ip, cp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_char_p
        L.dem_write_particle_state.argtypes = [cp, C.c_int, ip, ip, dp, dp, dp, dp, dp, dp]
        L.dem_write_particle_force.argtypes = [cp, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.dem_write_body_force.argtypes = [cp, C.c_int, C.c_int, dp]
        L.dem_write_body_state.argtypes = [cp, C.c_int, C.c_double, dp, dp, dp, dp, dp, dp]
        _io = L
    return _io


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _check(rc, path):
    if rc == -1:
        raise OSError(f"Cannot open output file: \"{os.path.abspath(path)}\"")
    if rc != 0:
        raise OSError(f"failed to write output file: \"{os.path.abspath(path)}\"")


def _prepare(path):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    return path


def write_particle_state(path, ids, status, mass, radius, pos, vel, quat, angvel):
    ids = np.ascontiguousarray(ids, np.int32)
    st = np.ascontiguousarray(status, np.int32)
    n = len(ids)
    _check(io_lib().dem_write_particle_state(
        _prepare(path).encode(), n, _p(ids, C.c_int32), _p(st, C.c_int32), _p(_d(mass)), _p(_d(radius)),
        _p(_d(pos, (n, 3))), _p(_d(vel, (n, 3))), _p(_d(quat, (n, 4))), _p(_d(angvel, (n, 3)))), path)


def write_particle_force(path, contact, wall, field, total, body_particle=None):
    n = len(contact)
    bp = None if body_particle is None else _d(body_particle)
    nb = 0 if bp is None else bp.shape[0]
    _check(io_lib().dem_write_particle_force(
        _prepare(path).encode(), n, nb, _p(_d(contact, (n, 6))), _p(_d(wall, (n, 6))), _p(_d(field)[:, :3].copy()),
        _p(_d(total, (n, 6))), _p(bp)), path)


def write_body_force(path, detail):
    d = _d(detail)
    nb, n = d.shape[0], d.shape[1]
    _check(io_lib().dem_write_body_force(_prepare(path).encode(), n, nb, _p(d)), path)


def write_body_state(path, time, center, vel, angvel, orien, loads, impulses=None):
    nb = len(center)
    _check(io_lib().dem_write_body_state(
        _prepare(path).encode(), nb, float(time), _p(_d(center, (nb, 3))), _p(_d(vel, (nb, 3))),
        _p(_d(angvel, (nb, 3))), _p(_d(orien, (nb, 4))), _p(_d(loads, (nb, 6))),
        None if impulses is None else _p(_d(impulses, (nb, 6)))), path)


def _g17(x):
    return "%.17g" % x


def regime_rows(stats, end_step, end_time):
    """writeBodyRegimeRows (cuda_simulation_integrator.cpp:90-108)."""
    ri, rv = stats["regime_instant"], stats["regime_interval"]
    out = []
    for b in range(len(ri)):
        cells = ["native_pt_branch_v1", str(int(end_step)), _g17(end_time), str(b),
                 str(int(stats["body_regime_interval_native_steps"]))]
        cells += [str(int(v)) for v in ri[b]] + [str(int(v)) for v in rv[b]]
        out.append(",".join(cells) + "\n")
    return "".join(out)


def micro_rows(stats, end_step, end_time):
    """writeBodyMicrostateRows (cuda_simulation_integrator.cpp:144-158)."""
    mi, mv = stats["micro_instant"], stats["micro_interval"]
    out = []
    for b in range(len(mi)):
        cells = ["native_pt_microstate_v1", str(int(end_step)), _g17(end_time), str(b),
                 str(int(stats["body_regime_interval_native_steps"]))]
        cells += [_g17(v) for v in mi[b]] + [_g17(v) for v in mv[b]]
        out.append(",".join(cells) + "\n")
    return "".join(out)


class CaseOutput:
    """Frame writer for one run (AsyncOutputWriter + the CSV streams of the
    coupled loop).  options: particle_state, particle_force, body_state,
    body_force (CONTROL_OUTPUT_*).  `frame` downloads from the solver on the
    calling thread; a worker thread formats and writes."""

    def __init__(self, case_dir, scene, particle_state=True, particle_force=True, body_state=True,
                 body_force=True, csv=None, depth=2):
        self.root = os.path.join(case_dir, "OutputFile")
        self.scene = scene
        self.nb = len(scene.body_mass)
        self.opt = dict(particle_state=particle_state, particle_force=particle_force,
                        body_state=body_state and self.nb > 0, body_force=body_force and self.nb > 0)
        self.csv = (self.nb > 0) if csv is None else csv
        self._ids = np.ascontiguousarray(scene.pid, np.int32)
        self._status = np.ascontiguousarray(scene.status, np.int32)
        self._mass = _d(scene.mass)
        self._radius = _d(scene.radius)
        self._q = queue.Queue(maxsize=depth)
        self._error = None
        self._thread = threading.Thread(target=self._work, daemon=True)
        self._thread.start()
        if self.csv:
            self._regime = open(_prepare(os.path.join(self.root, "native_body_contact_branches.csv")), "w")
            self._micro = open(_prepare(os.path.join(self.root, "native_body_contact_microstate.csv")), "w")
            self._regime.write(REGIME_HEADER)
            self._micro.write(MICRO_HEADER)

    def any(self):
        return any(self.opt.values())

    def _work(self):
        while True:
            job = self._q.get()
            if job is None:
                return
            try:
                if self._error is None:
                    job()
            except BaseException as e:  # reported on the next frame / finish
                self._error = e

    def _raise(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def frame(self, solver, step, time, extra=None):
        """enqueueCudaFrame (cuda_simulation_integrator.cpp:160-172)."""
        self._raise()
        if not self.any():
            return
        index = step + 1
        o = self.opt
        st = solver.state() if o["particle_state"] else None
        fo = None
        if o["particle_force"] or o["body_force"]:
            fo = solver.forces(body_particle=o["particle_force"] and self.nb > 0,
                               body_detail=o["body_force"])
        body = None
        if o["body_state"] and extra is not None:
            b = extra["bodies"]
            body = (time, b["center"].copy(), b["vel"].copy(), b["angvel"].copy(), b["orien"].copy(),
                    np.array(extra["loads"]), np.array(extra["impulses"]))
        root = self.root

        def job():
            if o["particle_force"]:
                write_particle_force(os.path.join(root, "force_particles", "F.%08d.bt" % index), fo["contact"],
                                     fo["wall"], fo["field"], fo["total"],
                                     fo["body_particle"] if self.nb else None)
            if o["body_force"]:
                write_body_force(os.path.join(root, "force_bodys", "FB.%05d.bt" % index), fo["body_detail"])
            if o["particle_state"]:
                write_particle_state(os.path.join(root, "state_particles", "ph.%08d.bt" % index), self._ids,
                                     self._status, self._mass, self._radius, st["pos"], st["vel"], st["quat"],
                                     st["angvel"])
            if body is not None:
                write_body_state(os.path.join(root, "state_bodys", "At.%08d.bt" % index), *body)

        self._q.put(job)

    def stats_rows(self, stats, end_step, end_time):
        if self.csv:
            self._regime.write(regime_rows(stats, end_step, end_time))
            self._micro.write(micro_rows(stats, end_step, end_time))

    def finish(self):
        self._q.put(None)
        self._thread.join()
        if self.csv:
            self._regime.close()
            self._micro.close()
        self._raise()
