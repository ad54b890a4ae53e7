"""Runs a reference case directory the way the reference's host does
(Simulation::run -> IntegrateDem / IntegrateDemMultiBody,
src/simulation.cpp:133-156, src/cuda_simulation_integrator.cpp:245-446) on
the HIP backend: reads <case>/InputFile/SettingData/world.par and the files
it names, restores <case>/InputFile/contact_history.bt when present, steps
ODE_StartStep .. ODE_EndStep, writes the frame files and CSVs selected by
CONTROL_OUTPUT_* under <case>/OutputFile/ and saves
<case>/OutputFile/contact_history.bt at the end.

    python -m dem_mbd_cuda_amd.run <case_dir>
"""
from __future__ import annotations

import os
import sys

from . import host, library, output, scene


def run_case(case_dir, lib=None, log=print):
    par = scene.read_world_par(os.path.join(case_dir, "InputFile", "SettingData", "world.par"))
    sc = scene.read_case(case_dir)
    flag = lambda key: int(par.get(key, "1")) != 0
    start = int(float(par.get("ODE_StartStep", "0")))
    keys = []
    if len(sc.body_mass):
        # read before the solver starts, as the reference does; the path is
        # opened as written, relative to the working directory
        # (cuda_simulation_integrator.cpp:182-185, 289-294)
        keys = host.load_prescribed_motion(par.get("CONTROL_PRESCRIBED_MOTION_FILE", ""), start,
                                           start + sc.steps)
        if keys:
            host.check_prescribed_bodies(sc)
    solver = (lib or library()).solver()
    solver.initialize(sc)
    out = output.CaseOutput(case_dir, sc, particle_state=flag("CONTROL_OUTPUT_PARTICLE_STATE"),
                            particle_force=flag("CONTROL_OUTPUT_PARTICLE_FORCE"),
                            body_state=flag("CONTROL_OUTPUT_BODY_STATE"),
                            body_force=flag("CONTROL_OUTPUT_BODY_FORCE"))
    hin = os.path.join(case_dir, "InputFile", "contact_history.bt")
    hout = os.path.join(case_dir, "OutputFile", "contact_history.bt")
    os.makedirs(os.path.dirname(hout), exist_ok=True)
    dt = sc.dt
    try:
        if len(sc.body_mass):
            def frame(step, t, extra):
                out.frame(solver, step, t, extra)
                if step < 0:
                    out.stats_rows(solver.stats(), start, start * dt)
                else:
                    out.stats_rows(solver.stats(), step + 1, (step + 1) * dt)
                    log(f"  {step + 1}/{start + sc.steps}")

            host.run_coupled(solver, sc, start_step=start, on_frame=frame, history_in=hin, history_out=hout,
                             prescribed=keys)
        else:
            def frame(step, t, extra):
                out.frame(solver, step, t, extra)
                if step >= 0:
                    log(f"Step = {step + 1}/{start + sc.steps}, contacts = {solver.stats()['particle_contacts']}")

            host.run_pure(solver, sc, start_step=start, on_frame=frame, history_in=hin, history_out=hout)
    finally:
        out.finish()
        solver.close()


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m dem_mbd_cuda_amd.run <case_dir>", file=sys.stderr)
        return 2
    run_case(argv[0])
    return 0


if __name__ == "__main__":
    sys.exit(main())
