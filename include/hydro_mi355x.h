/*
 * hydro_mi355x.h -- entry points this build adds next to the reference API
 * (hydro_solver_3d.h, gravity_solver_3d.h, input_deck.h).
 *
 * The reference keeps its step function static (hydro_solver_3d.c:27,
 * :637) and its state in host memory; these functions expose the device-
 * resident step, explicit host<->device synchronisation points, the
 * extended Poisson settings and diagnostics.  Every function returns 0 on
 * success and a negative value on error (a message goes to stderr); none
 * aborts the process.
 */
#ifndef HYDRO_MI355X_H
#define HYDRO_MI355X_H

#include "gravity_solver_3d.h"
#include "hydro_solver_3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device and state transfer ---------------------------------------- */

/* Number of visible GPUs (0 when none); never initialises a context on
 * failure paths. */
int hydro_device_count(void);

/* Copy the host mirror solver->U (AoS, 40 B per cell) to the device and
 * restart the dt pipeline from it.  The run loops call this themselves at
 * start; callers that write solver->U between steps call it explicitly. */
int solver3d_upload(GodunovHLLC3D *solver);
/* Copy the device state to solver->U and the gravitational acceleration in
 * use to solver->gx/gy/gz. */
int solver3d_download(GodunovHLLC3D *solver);
/* Replace the state: U_aos (nx*ny*nz*5 doubles) into solver->U, then upload. */
int solver3d_set_state(GodunovHLLC3D *solver, const double *U_aos);

/* ---- stepping ---------------------------------------------------------- */

/* One step of the reference step function (godunov_step_3d,
 * hydro_solver_3d.c:637-811) on the device: gravity solve when enabled, dt,
 * fluxes, update, gravity source, boundaries.  Sets solver->dt. */
int solver3d_step(GodunovHLLC3D *solver);
/* nsteps steps with no host synchronisation between them when gravity is
 * off.  dt_log (may be NULL) receives the dt of each step.  solver->dt is
 * the last step's dt on return. */
int solver3d_run_steps(GodunovHLLC3D *solver, int nsteps, double *dt_log);
/* The reference time loop without output (hydro_solver_3d.c:2257-2260):
 * while (t < t_end) { step; t += dt; }, stopping early after max_steps
 * (<= 0: no limit).  Logs (may be NULL, length >= log_cap) receive dt and
 * SOR iterations per step.  Returns the number of steps (< 0 on error). */
int solver3d_run(GodunovHLLC3D *solver, double t_end, int max_steps, double *t_out,
                 double *dt_log, int *iter_log, int log_cap);
/* Run nsteps steps and time them on the library's stream with HIP events:
 * wall_ms = whole loop (gravity included), step_ms / bc_ms = average
 * duration of the fused step kernel and of the boundary kernel. */
int solver3d_benchmark_steps(GodunovHLLC3D *solver, int nsteps, double *wall_ms,
                             double *step_ms, double *bc_ms);
/* Device milliseconds of the last hydro kernels (step + boundary) and of
 * the last gravity solve. */
int solver3d_last_timing(const GodunovHLLC3D *solver, double *hydro_ms, double *gravity_ms);

/* Diagnostic: number of the n quotients a[i] / b[i] (host arrays) for which
 * the reciprocal-corrected division used by the step kernels differs in any
 * bit from the IEEE division on the device (0 expected), < 0 on error. */
long hydro_check_division(const double *a, const double *b, long n);

/* ---- gravity settings (no reference counterpart) ----------------------- */

/* Multipole order of the isolated boundary: 2 = monopole + quadrupole as
 * the reference (gravity_solver_3d.c:302-497); up to 8. */
int gravity_set_multipole_lmax(GravitySolver3D *solver, int lmax);
/* POISSON_FP64_SOR or POISSON_FP32_DC (input_deck.h). */
int gravity_set_poisson_mode(GravitySolver3D *solver, int mode);
/* SOR_CRIT_MAX_UPDATE (the reference) or SOR_CRIT_REL_RESID. */
int gravity_set_criterion(GravitySolver3D *solver, int criterion);
/* Make the device potential equal to solver->phi (a caller-provided
 * initial guess), or copy the device potential and g back to the host. */
int gravity_sync_to_device(GravitySolver3D *solver);
int gravity_sync_to_host(GravitySolver3D *solver);
/* Moments of a host density as the boundary uses them: out13 = M, xcm,
 * ycm, zcm, Q[3][3]; out_hi (may be NULL) = A_lm, B_lm for l = 3..lmax. */
int gravity_moments(GravitySolver3D *solver, const double *rho, double *out13, double *out_hi);
/* Boundary potential of a host density written into phi_out (all cells
 * copied from the device potential, boundary faces replaced). */
int gravity_boundary_potential(GravitySolver3D *solver, const double *rho, double *phi_out);
/* ||4 pi G rho - L_h phi||_2 / ||4 pi G rho||_2 of the device potential for
 * a host density (interior cells). */
int gravity_rel_residual(GravitySolver3D *solver, const double *rho, double *rel);

/* Poisson solve on a device density (nx*ny*nz doubles in device memory);
 * the potential and (when converged) g stay on the device. */
int gravity_solve_device(GravitySolver3D *solver, const double *d_rho);

#ifdef __cplusplus
}
#endif
#endif /* HYDRO_MI355X_H */
