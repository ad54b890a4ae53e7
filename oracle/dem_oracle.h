/*
 * dem_oracle.h -- CPU restatement of the reference's FP64 DEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * in libdemhip.so: only the tests, the smoke check and the CPU-baseline leg
 * of bench.py load it.  No product code links or calls it.
 *
 * It restates, in plain C with a straightforward loop structure, the
 * algorithm of the reference's CUDA backend (src/cuda/cuda_contact_backend.cu,
 * cuda_grid_backend.cu, cuda_particle_backend.cu, cuda_body_backend.cu and the
 * sequencing in src/cuda_backend_cuda.cu).  Each function in dem_oracle.c
 * cites the reference lines it follows.  The API mirrors include/dem_hip.h
 * entry for entry (ora_ prefix) so a test can drive both with one script.
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, IEEE double), which is
 * also how the HIP kernels are compiled, so both evaluate every expression in
 * the same order with correctly rounded +,-,*,/,sqrt.
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#include "../include/dem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_solver ora_solver;

int ora_create(ora_solver** out);
void ora_destroy(ora_solver* s);
/* Threads of the per-owner contact loop (OpenMP; default: all the runtime
 * offers).  Results do not depend on it.  No reference counterpart. */
int ora_set_threads(ora_solver* s, int threads);
int ora_initialize(ora_solver* s, const dem_particles_view* particles,
                   const dem_components_view* components,
                   const dem_walls_view* walls, const dem_bodies_view* bodies,
                   const dem_mech* pp, const dem_mech* pt, const dem_mech* pw,
                   const dem_control* control);
int ora_compute_forces(ora_solver* s, const double* body_pose13,
                       double* body_loads6, int collect);
int ora_init_half_step(ora_solver* s, double half_dt);
int ora_advance_particles(ora_solver* s, double full_dt, double* vmax_out);
int ora_finish_particle_step(ora_solver* s, double full_dt, double half_dt);
int ora_download_particle_state(const ora_solver* s, double* pos, double* vel,
                                double* ang_vel, double* quat);
int ora_download_force_output(const ora_solver* s, double* contact,
                              double* field, double* wall, double* total,
                              double* body_particle, double* body_detail);
int ora_get_step_stats(const ora_solver* s, dem_step_stats* stats,
                   uint64_t* regime_instant, uint64_t* regime_interval,
                   double* micro_instant, double* micro_interval);
int ora_particle_grid_entry_count(const ora_solver* s, size_t* out);
int ora_triangle_grid_entry_count(const ora_solver* s, size_t* out);
int ora_debug_grid(const ora_solver* s, uint64_t* sorted_keys,
                   int32_t* sorted_ids);
int ora_debug_history(const ora_solver* s, uint64_t* pp_keys,
                      double* pp_values, uint64_t* pw_keys, double* pw_values,
                      uint64_t* pt_keys, double* pt_values);
int ora_debug_triangle_grid(const ora_solver* s, uint64_t* keys,
                            int32_t* triangles);
int ora_save_contact_history(const ora_solver* s, const char* path);
int ora_load_contact_history(ora_solver* s, const char* path);
int ora_has_restored_leapfrog_state(const ora_solver* s, int* out);
const char* ora_last_error(void);

/* Geometry kernels exposed for the known-answer tests. */
double ora_circle_triangle_area(const double center[3], const double normal[3],
                                const double a[3], const double b[3],
                                const double c[3], double radius);
uint64_t ora_pack_cell(int x, int y, int z);
double ora_atan2(double y, double x);

#ifdef __cplusplus
}
#endif

#endif
