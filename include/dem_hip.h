/*
 * dem_hip.h -- C-ABI of libdemhip.so, the MI355X (gfx950) FP64 DEM step.
 *
 * This is the drop-in boundary for the reference's persistent GPU solver
 * (`class CudaDemSolver` + `CudaBackendAvailable` + `CudaDeviceDescription`,
 * reference include/cuda_backend.h:41-75).  Every entry point below replaces
 * one member of that class; the comment on each names the member (file:line)
 * and states where the argument meaning differs (plain pointers and sizes
 * instead of the host's PARTICLE/BODYSET/FORCE classes).
 *
 * Conventions
 *   - Every function returns DEM_OK (0) or a negative DEM_E* code.  On error
 *     dem_last_error() returns the message of the failing call on this host
 *     thread.  The text matches the std::runtime_error text the reference
 *     throws for the same condition, so a C++ adapter can rethrow it as is.
 *   - All arrays are host memory owned by the caller, row-major AoS as in
 *     the host classes: vec3 = 3 doubles, quaternion = (w, x, y, z).
 *   - Results are complete in the caller's buffers when a call returns.
 *   - One solver = one device (device 0) and one HIP stream.  A solver is
 *     not thread-safe; distinct solvers may be used from distinct threads.
 */
#ifndef DEM_HIP_H
#define DEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_OK 0
#define DEM_E_INVALID (-1)   /* bad argument / shape / state           */
#define DEM_E_DEVICE (-2)    /* HIP runtime error                      */
#define DEM_E_CAPACITY (-3)  /* history / grid / index capacity        */
#define DEM_E_IO (-4)        /* checkpoint file error                  */
#define DEM_E_UNSUPPORTED (-5)

/* History capacities per component (reference cuda_contact_backend.cuh:11-13). */
#define DEM_PARTICLE_HISTORY_SLOTS 64
#define DEM_WALL_HISTORY_SLOTS 16
#define DEM_BODY_HISTORY_SLOTS 64
#define DEM_BODY_REGIME_COUNTERS 7      /* cuda_contact_backend.cuh:15 */
#define DEM_BODY_MICROSTATE_VALUES 22   /* cuda_contact_backend.cuh:16 */

typedef struct dem_solver dem_solver;

/* Owner particles (host PARTICLE, reference include/particle.h:5-20). */
typedef struct dem_particles_view {
    int32_t count;            /* PARTICLE::Num                           */
    const int32_t* id;        /* [n] stable ids, PARTICLE::Number        */
    const int32_t* status;    /* [n] PARTICLE::Status (-1 = frozen)      */
    const double* mass;       /* [n]                                     */
    const double* radius;     /* [n]                                     */
    const double* inertia;    /* [n][3] body-frame principal moments     */
    const double* pos;        /* [n][3] world                            */
    const double* vel;        /* [n][3] world                            */
    const double* ang_vel;    /* [n][3] body frame                       */
    const double* quat;       /* [n][4] (w, x, y, z)                     */
} dem_particles_view;

/* Sphere components of each owner (PARTICLE::Component*, particle.h:9-13). */
typedef struct dem_components_view {
    int32_t count;            /* PARTICLE::ComponentNum                  */
    const int32_t* id;        /* [c] ComponentId (>= 0)                  */
    const int32_t* owner;     /* [c] zero-based owner row                */
    const double* radius;     /* [c]                                     */
    const double* pos_body;   /* [c][3] offset in the owner body frame   */
} dem_components_view;

/* Infinite planar walls (WALL::Orig / WALL::N, include/wall.h:5-9). */
typedef struct dem_walls_view {
    int32_t count;
    const double* origin;     /* [w][3]                                  */
    const double* normal;     /* [w][3] unit, pointing into the domain   */
} dem_walls_view;

/* Rigid triangle-mesh bodies (BODYSET/BODY/STLTetMesh, multibody.h:28-87).
 * pose13 per body = centre (3), orientation quaternion (4), linear velocity
 * (3), angular velocity in the body frame (3): the 13 doubles the reference
 * uploads each step (cuda_backend_cuda.cu:205-227). */
typedef struct dem_bodies_view {
    int32_t count;            /* BODYSET::Num (0 = pure DEM)             */
    const double* mass;       /* [b]  BODY::Mass (host adds m*g)         */
    const double* pose13;     /* [b][13]                                 */
    const int32_t* tri_count; /* [b]  BODY::Size                         */
    const double* tri_local;  /* [sum tri][15]: v1, v2, v3, v4, normal   */
} dem_bodies_view;

/* One MECH block (include/simulation.h:48-64) as the kernels use it. */
typedef struct dem_mech {
    double epsS, mu, epsN, cohesion, beta, muT, muR, kN;
} dem_mech;

/* CONTROL + ODE fields read by the solver (simulation.h:10-46). */
typedef struct dem_control {
    double dt;                    /* ODE::StepSize                       */
    double gravity[3];            /* CONTROL::g                          */
    int32_t gravity_enabled;      /* CONTROL::Grav_flag != 0             */
    int32_t universal_gravity;    /* CONTROL::Universal_gravitation_flag */
    int32_t keep_body_particle;   /* CONTROL::OutputParticleForce        */
    int32_t keep_body_detail;     /* CONTROL::OutputBodyForce            */
    double mesh_scale;            /* cell = mesh_scale * max component
                                     radius; 0 selects the reference's 2.5
                                     (cuda_backend_cuda.cu:378)          */
} dem_control;

/* CudaStepStats scalars (include/cuda_backend.h:17-27). */
typedef struct dem_step_stats {
    int32_t particle_contacts;
    int32_t body_contacts;
    double body_max_depth;
    double body_contact_fz;
    int32_t particle_history_overflows;
    int32_t body_history_overflows;
    int32_t particle_history_high_water;
    int32_t body_history_high_water;
    int32_t triangle_grid_overflows;
    int32_t reserved;
    uint64_t body_regime_interval_native_steps;
} dem_step_stats;

/* ---- solver lifetime: CudaDemSolver(), ~CudaDemSolver() (cuda_backend.h:45-46) */
int dem_create(dem_solver** out);
void dem_destroy(dem_solver* solver);

/* CudaDemSolver::initialize (cuda_backend.h:50-53; cuda_backend_cuda.cu:302-381).
 * walls/bodies may be NULL or have count 0.  pp/pt/pw are the particle-
 * particle, particle-triangle and particle-wall MECH blocks.  Fails with the
 * reference's messages on re-initialisation, duplicate particle ids, bad
 * component topology (1..8 components per owner, ids >= 0, radius > 0).
 * Owners with several components (clumps) are stepped with the contact
 * grid, scan and laws over the components, each owner summing its
 * components in (owner, component id) order. */
int dem_initialize(dem_solver* solver, const dem_particles_view* particles,
                   const dem_components_view* components,
                   const dem_walls_view* walls, const dem_bodies_view* bodies,
                   const dem_mech* pp, const dem_mech* pt, const dem_mech* pw,
                   const dem_control* control);

/* A batch of independent scenes stepped together by one solver.  No
 * reference counterpart: the reference runs one scene per process
 * (cuda_simulation_integrator.cpp:248-285 drives one CudaDemSolver); a
 * parameter sweep over many small scenes is launch-bound that way.  The
 * particle, component and wall views hold the scenes back to back; scene k
 * owns particle rows [particle_start[k], particle_start[k+1]) and walls
 * [wall_start[k], wall_start[k+1]), has its own particle-particle and
 * particle-wall MECH blocks and its own grid cell (mesh_scale times the
 * scene's largest radius).  Particles interact only within their scene, and
 * every per-scene result equals that of a solver initialised with the scene
 * alone (bit for bit).  Particle ids must be unique within a scene.  Wall
 * history keys are the wall's index within its scene + 1.  A batch holds
 * single-sphere owners only (clumps: DEM_E_UNSUPPORTED). */
typedef struct dem_batch_view {
    int32_t scene_count;
    const int64_t* particle_start; /* [scene_count + 1], 0 .. particle count */
    const int32_t* wall_start;     /* [scene_count + 1], 0 .. wall count      */
    const dem_mech* pp;            /* [scene_count] particle-particle         */
    const dem_mech* pw;            /* [scene_count] particle-wall             */
} dem_batch_view;

/* Initialises a solver with a batch of scenes (no bodies, no universal
 * gravitation); afterwards every other entry point applies to the whole
 * batch (downloads return the scenes back to back, statistics are totals). */
int dem_initialize_batch(dem_solver* solver, const dem_particles_view* particles,
                         const dem_components_view* components,
                         const dem_walls_view* walls, const dem_batch_view* batch,
                         const dem_control* control);

/* CudaDemSolver::computeForces (cuda_backend.h:54-55; cuda_backend_cuda.cu:383-524).
 * body_pose13: [b][13] current body poses or NULL (no body upload, as when the
 * host passes a null BODYSET).  body_loads6: [b][6] out, contact load on each
 * body plus m_b*g when gravity is on (the host's double** bodyForces rows), or
 * NULL.  collect != 0 downloads the step statistics and raises the history /
 * triangle-grid capacity errors, as the reference does on diagnostics steps. */
int dem_compute_forces(dem_solver* solver, const double* body_pose13,
                       double* body_loads6, int collect);

/* No reference counterpart.  dem_compute_forces for callers that cannot
 * allocate device memory (serving loops and harnesses that forbid hipMalloc,
 * code that runs between graph captures).  Same arguments and results, bit
 * for bit.  The only device memory dem_compute_forces allocates after
 * initialisation is the triangle grid's per-step buffers (sorted entries,
 * cell table, sort counts), grown with hipMalloc/hipFree when a step needs
 * more; here they come from one static device array of
 * DEM_NOALLOC_SCRATCH_BYTES that every device the library runs on holds for
 * the life of the process, and the call makes no hipMalloc or hipFree.  A step
 * whose triangle grid needs more than the array fails with DEM_E_CAPACITY
 * before any kernel touches it.  Solvers without triangles use no scratch.
 * The call returns after its kernels are done with the array.  All calls on
 * one device share its array: calls that overlap in time (solvers driven from
 * several host threads) race, and dem_debug_triangle_grid after such a call
 * reads the array as the device's last no-allocation call left it. */
#define DEM_NOALLOC_SCRATCH_BYTES (32u << 20)
int dem_compute_forces_noalloc(dem_solver* solver, const double* body_pose13,
                               double* body_loads6, int collect);

/* CudaDemSolver::initializeParticleHalfStep (cuda_backend.h:56). */
int dem_init_half_step(dem_solver* solver, double half_dt);
/* CudaDemSolver::advanceParticles (cuda_backend.h:57); *vmax_out gets 0.0
 * like the reference (cuda_backend_cuda.cu:531-538); may be NULL. */
int dem_advance_particles(dem_solver* solver, double full_dt, double* vmax_out);
/* CudaDemSolver::finishParticleStep (cuda_backend.h:58). */
int dem_finish_particle_step(dem_solver* solver, double full_dt, double half_dt);

/* CudaDemSolver::loadContactHistory / saveContactHistory /
 * hasRestoredParticleLeapfrogState (cuda_backend.h:59-61): the DMHIST2
 * checkpoint of cuda_backend_cuda.cu:545-822. */
int dem_load_contact_history(dem_solver* solver, const char* path);
int dem_save_contact_history(const dem_solver* solver, const char* path);
int dem_has_restored_leapfrog_state(const dem_solver* solver, int* out);

/* CudaDemSolver::downloadParticleState (cuda_backend.h:63): owner state in
 * input row order; any pointer may be NULL to skip that field. */
int dem_download_particle_state(const dem_solver* solver, double* pos,
                                double* vel, double* ang_vel, double* quat);

/* CudaDemSolver::downloadForceOutput (cuda_backend.h:64-65): FORCE rows.
 * contact/field/wall/total: [n][6] (Fc, Ff, Fw, FR; the host sets Fd = 0).
 * body_particle: [b][n][6] (FI_pt), needs keep_body_particle.
 * body_detail:   [b][n][19] (ForceBody), needs keep_body_detail.
 * Any pointer may be NULL. */
int dem_download_force_output(const dem_solver* solver, double* contact,
                              double* field, double* wall, double* total,
                              double* body_particle, double* body_detail);

/* CudaDemSolver::stepStats (cuda_backend.h:66): values cached by the last
 * compute call with collect != 0.  The per-body arrays are optional
 * ([b][7] u64 counters and [b][22] f64 microstate sums). */
int dem_get_step_stats(const dem_solver* solver, dem_step_stats* stats,
                       uint64_t* regime_instant, uint64_t* regime_interval,
                   double* micro_instant, double* micro_interval);

/* particleGridEntryCount / triangleGridEntryCount / deviceMemoryBytes
 * (cuda_backend.h:67-69). */
int dem_particle_grid_entry_count(const dem_solver* solver, size_t* out);
int dem_triangle_grid_entry_count(const dem_solver* solver, size_t* out);
int dem_device_memory_bytes(const dem_solver* solver, size_t* out);

/* CudaBackendAvailable / CudaDeviceDescription (cuda_backend.h:74-75). */
int dem_backend_available(void);
const char* dem_device_description(void);

/* Thread-local message of the last failed call ("" if none). */
const char* dem_last_error(void);

/* ---- introspection for parity checks (no reference counterpart) ----------
 * Reads the state left by the last dem_compute_forces, in the reference's
 * own index space (components sorted by (owner, component id)).
 * dem_debug_grid: the particle grid as the reference sorts it: 64-bit packed
 * cell keys and component indices in (key, index) order, [c] each; after an
 * evaluation that sorted the grid this is the device's sorted order itself.
 * dem_debug_history: history tables in the reference's [component][slot]
 * layout; keys are the reference's (other id + 1, wall + 1, body + 1), 0 for
 * an empty slot; values are [component][slot][9] (roll, tangent, twist).
 * Any pointer may be NULL. */
int dem_debug_grid(const dem_solver* solver, uint64_t* sorted_keys,
                   int32_t* sorted_ids);
int dem_debug_history(const dem_solver* solver, uint64_t* pp_keys,
                      double* pp_values, uint64_t* pw_keys, double* pw_values,
                      uint64_t* pt_keys, double* pt_values);
/* Counters of the solver's work since initialisation (no reference
 * counterpart; for measurement and tests), out[0..8] as far as capacity
 * allows: force evaluations; evaluations that sorted the particle grid on
 * the device; candidate-list rebuilds; steps of dem_run_steps replayed from
 * HIP graphs; steps of dem_run_steps launched eagerly; 1 if the last
 * evaluation sorted the grid (dem_debug_grid then returns the device's own
 * sorted order as it is; between candidate-list rebuilds it puts the current
 * cells in (key, index) order on the host); block passes of the full 27-cell
 * scan on diagnostics evaluations whose candidates were staged in LDS, and
 * those that read them from device memory; the candidate-list skin in
 * picometres (0 without candidate lists). */
int dem_debug_run_counters(const dem_solver* solver, long long* out, int capacity,
                           int* count);
/* Triangle grid entries (key, triangle) in sorted order, [entries] each. */
int dem_debug_triangle_grid(const dem_solver* solver, uint64_t* keys,
                            int32_t* triangles);
/* Blocks until all work queued on the solver's stream has finished (no
 * reference counterpart: the reference's calls are synchronous, these are
 * ordered on one stream and complete at the next download). */
int dem_synchronize(const dem_solver* solver);
/* Turns on (enabled != 0) or off the stage events behind
 * dem_debug_last_timings; off by default, since each event adds a short gap
 * between kernels (no reference counterpart). */
int dem_debug_enable_timings(dem_solver* solver, int enabled);
/* Device kernel timing of the last dem_compute_forces made while timings
 * were enabled (count 0 otherwise), in milliseconds:
 * [0] grid build, [1] contact scan + list compaction, [2] per-contact
 * particle law, [3] per-particle sums (triangles, walls, field,
 * acceleration), [4] body reductions.  For measurement only. */
int dem_debug_last_timings(const dem_solver* solver, double* ms_out,
                           int capacity, int* count);

/* ---- multi-step runs (no reference counterpart) --------------------------
 * Runs `steps` iterations of the reference's pure-DEM loop body
 * (advanceParticles, computeForces without diagnostics, finishParticleStep;
 * cuda_simulation_integrator.cpp:263-269), identical in every result to
 * calling those three entry points in turn.  Without bodies the steps are
 * captured into HIP graphs and replayed (DEM_RUN_NO_GRAPH: launch eagerly);
 * DEM_RUN_AUTOTUNE times the reorder interval and the per-contact grid on
 * the run's first steps (once per solver) and keeps the fastest.  The last
 * step's finish is left pending, as after dem_finish_particle_step.
 * *steps_done (may be NULL) receives the steps run. */
#define DEM_RUN_NO_GRAPH 1
#define DEM_RUN_AUTOTUNE 2
int dem_run_steps(dem_solver* solver, int steps, double dt, int flags,
                  int* steps_done);

/* ---- differentiable runs (no reference counterpart) ---------------------
 * Reverse-mode derivatives of scene observables with respect to each
 * scene's contact laws, for a solver freshly initialised (dem_initialize or
 * dem_initialize_batch) with spheres, walls and gravity only.  The run is the
 * reference's pure-DEM loop (cuda_simulation_integrator.cpp:245-286): the
 * first force evaluation, the half-step initialisation, then `steps` x
 * (advance, forces, finish), without particle reordering.  Per scene it
 * reports the mean centre height z after the last step and the plate load:
 * the z force of wall `plate_wall` (an index within each scene's walls) on
 * the scene's particles, averaged over the `steps` force evaluations after
 * the first.  The forward run keeps a checkpoint of the particle state and
 * contact history every `checkpoint_interval` steps on the device. */
int dem_diff_forward(dem_solver* solver, int steps, int checkpoint_interval,
                     int plate_wall, double* bed_height /* [scenes] */,
                     double* plate_load /* [scenes] */);
/* After dem_diff_forward: given d(objective)/d(bed_height) and
 * d(objective)/d(plate_load) per scene, returns d(objective)/d(MECH) per
 * scene as [scene][8] = {mu, epsN, epsS, kN} of the particle-particle law,
 * then the same four of the particle-wall law.  Recomputes every checkpoint
 * interval from its checkpoint and sweeps it backwards; may be called again
 * with other seeds.  Leaves the solver at the end of the run. */
int dem_diff_backward(dem_solver* solver, const double* d_bed_height,
                      const double* d_plate_load, double* d_mech /* [scenes][8] */);

/* Host evaluations of one particle-particle / particle-wall contact law and
 * of its reverse-mode derivative (the arithmetic the kernels run), for
 * derivative checks; the packed layouts are documented in
 * dem_mbd_cuda_amd/csrc/dem_diff_checks.inl.  No device is needed. */
int dem_check_pp_law(const double* in, double* out);
int dem_check_pp_law_vjp(const double* in, const double* seeds, double* grads);
int dem_check_pw_law(const double* in, double* out);
int dem_check_pw_law_vjp(const double* in, const double* seeds, double* grads);

#ifdef __cplusplus
}
#endif

#endif /* DEM_HIP_H */
