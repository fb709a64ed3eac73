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

/* ---- checkpoints (no reference counterpart: the reference restarts from
 * its %.6e CSV output, hydro_solver_3d.c:2139-2235, which drops the
 * potential, so SOR starts cold, gravity_solver_3d.c:87) ------------------ */

/* Full-precision checkpoint "HYDCK001": a 40-byte header (magic, int32 nx
 * ny nz, int32 flags: 1 potential, 2 g; double t; int32 output_count; int32
 * 0), U as the reference's AoS doubles, then with gravity the potential in
 * natural order and the gx, gy, gz planes.  Written from the device state. */
int solver3d_write_checkpoint(GodunovHLLC3D *solver, const char *path, double t,
                              int output_count);
/* Load a checkpoint (or a U-only "HYDIC001" initial-state file) into the
 * solver: U, and with gravity the potential the next solve warm-starts from
 * and the g a step uses when its solve does not converge, so a run resumed
 * from it continues bit for bit.  t and output_count (may be NULL) receive
 * the header's values (0 for an initial-state file).  The solver is left
 * unchanged on error (-1).  load_from_csv_restart and the deck's ic_file
 * accept the same files. */
int solver3d_read_checkpoint(GodunovHLLC3D *solver, const char *path, double *t,
                             int *output_count);

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

/* The reference's initial-condition problems into a host state array U
 * (nx*ny*nz cells, the reference's layout) without a solver or a device:
 * with deck == NULL as set_initial_conditions_3d(problem)
 * (hydro_solver_3d.c:1383-1515), otherwise as
 * set_initial_conditions_3d_from_deck(deck) (hydro_solver_3d.c:2092-2133).
 * Profile files are read from the paths the reference reads.  Returns 0, 1
 * for a problem type with no generator (U untouched), -1 on bad arguments. */
int hydro_fill_initial_state(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                             double ymax, double zmin, double zmax, double gamma, double G,
                             int problem, const InputDeck *deck, State3D *U);

/* Diagnostic: the sum ((0 + x[0]) + x[1]) + ... of n doubles (host array)
 * with every addition rounded in turn, as the reference's moment loops add
 * (gravity_solver_3d.c:182-198), computed on the device by method 0 (one
 * chain of additions), 1 (chunks that start at the running sum's predicted
 * peaks, each run from one start value per residue of the start's low bits
 * with a certificate of how far a start may move, and the exact sum carried
 * through those tables group by group), 2 (method 1, then every chunk run
 * again from its exact start; any mismatch redoes the sum by the chain) or 3
 * (method 1 with the tables built by the LDS-staged kernel that the moment
 * sums use, which a single sum would otherwise not reach).
 * stats (9 values, or NULL) receives methods 1 and 2's counters: chunks,
 * chunks with 1-3 and with 4-6 residue bits, chunks added term by term,
 * chunks whose re-run missed (method 2), clock cycles of the walk from group
 * to group, groups it walked chunk by chunk, 1 when the chain redid the sum,
 * chunks of zeros passed through.  *status = 0 on success. */
double hydro_sequential_sum(const double *x, long n, int method, long long *stats, int *status);

/* ---- gravity settings (no reference counterpart) ----------------------- */

/* Multipole order of the isolated boundary: 2 = monopole + quadrupole as
 * the reference (gravity_solver_3d.c:302-497); up to 8. */
int gravity_set_multipole_lmax(GravitySolver3D *solver, int lmax);
/* POISSON_FP64_SOR, POISSON_FP32_DC or POISSON_NN_FP32_DC (input_deck.h). */
int gravity_set_poisson_mode(GravitySolver3D *solver, int mode);
/* DC_INNER_SOR or DC_INNER_MULTIGRID (input_deck.h) for the FP32 modes;
 * multigrid needs even nx, ny, nz >= 8 (returns -1 otherwise). */
int gravity_set_dc_inner(GravitySolver3D *solver, int inner);
/* MOMENT_ORDER_TREE, MOMENT_ORDER_SEQUENTIAL or MOMENT_ORDER_AUTO (the
 * default; input_deck.h): how the isolated boundary's moment sums are
 * added.  The sequential order stores every term (max(4, nq) doubles per
 * cell, nq = 9 + sum_{l=3..lmax} 2(l+1)) and adds them one after another,
 * as compute_mass_moments and compute_quadrupole_moment do
 * (gravity_solver_3d.c:164-262), giving the reference's bits; AUTO takes it
 * for plain FP64 SOR and the compensated tree for the FP32 modes. */
int gravity_set_moment_order(GravitySolver3D *solver, int order);
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

/* ---- no device allocation (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc) ------------------------- */

/* Largest scratch (doubles) gravity_moments_noalloc holds: 64 MiB per GPU,
 * kept for the life of the process; enough for 256^3 at lmax <= 6 and for
 * 128^3 at any lmax. */
#define GRAVITY_MOMENTS_NOALLOC_MAX_SCRATCH (8L << 20)
/* Scratch doubles a grid needs (-1 for bad arguments); the call below
 * rejects grids needing more than the maximum above. */
long gravity_moments_noalloc_scratch(int nx, int ny, int nz, int lmax);
/* The moments gravity_moments returns for MOMENT_ORDER_TREE (same kernels,
 * same bits) of a device density d_rho (nx*ny*nz doubles) into the device
 * array d_mom: M, xcm, ycm, zcm, Q[3][3], then A_lm, B_lm for l = 3..lmax
 * (1 + (lmax+1)(lmax+2) doubles).  The grid spacing is (max - min)/n as in
 * gravity_solver_create.  Allocates nothing: the partial sums live in a
 * static device array shared by every call on that GPU, so calls that
 * overlap on one GPU race.  Synchronous; returns 0, -1 for bad arguments or
 * a grid beyond the scratch maximum, -2 on a device error. */
int gravity_moments_noalloc(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                            double ymax, double zmin, double zmax, int lmax,
                            const double *d_rho, double *d_mom);

/* ---- learned initial guess of the Poisson solve (no reference counterpart;
 * the reference starts SOR from the previous potential or from zero,
 * gravity_solver_3d.c:87 and :566) ---------------------------------------- */

#define NN_MAX_BLOBS 64
#define NN_MAX_MODES 256

/* Network with He-initialised weights from `seed`; NULL on failure. */
void *nn_model_create(unsigned long long seed);
void nn_model_destroy(void *model);
/* Number of trainable parameters, and the length of the packed parameter
 * vector (weights [27][cin padded to 4][cout] then bias, per layer). */
long nn_model_num_params(void *model);
long nn_model_param_len(void *model);
long nn_model_param_offset(void *model, int layer, int bias);
int nn_model_save(void *model, const char *path);
/* Loads weights and the network's kind and C (HYDNN002 files: kind 0). */
int nn_model_load(void *model, const char *path);
/* Kind 0: the potential guess (inputs from the density).  Kind 1: the start
 * correction; input the residual r = f - L phi of a start averaged onto the
 * network's grid (features r / s, -r / s, s its root mean square), output the
 * correction in units of C L^2 s (L the box edge, C = cscale > 0). */
int nn_model_set_kind(void *model, int kind, double cscale);
int nn_model_kind(void *model, double *cscale);
int nn_model_get_params(void *model, float *out);
int nn_model_set_params(void *model, const float *in);
/* One forward + backward pass on host data (kernel self-check). */
int nn_model_forward_backward(void *model, const float *rho, const float *target, int n,
                              const double *mom4, double L, float *loss, float *out,
                              float *feat, float *grads);
/* Device-resident training set of `count` n^3 fields. */
void *nn_dataset_create(int n, int count);
void nn_dataset_destroy(void *dataset);
/* Field `index`: blobs (cx cy cz mass width type, type 0 Gaussian, 1 n=1
 * polytrope of radius 2 width) on floor_rho, times exp(delta - sigma^2/2)
 * with delta = sum of amp cos(2 pi k.x + phase) over modes (kx ky kz phase
 * amp); solved cold by `grav` (an isolated n^3 solver).  Returns its status. */
int nn_dataset_fill(void *dataset, int index, const double *blob, int nblob, const double *mode,
                    int nmode, double floor_rho, double sigma, GravitySolver3D *grav);
int nn_dataset_get(void *dataset, int index, float *rho, float *phi, double *mom4);
/* Field `index` from host arrays: the network's input (density for kind 0,
 * averaged residual for kind 1) and its target (n^3 FP32 each). */
int nn_dataset_set(void *dataset, int index, const float *input, const float *target);
/* Adam over mini-batches of dataset fields listed in order[steps * batch]. */
int nn_train(void *model, void *dataset, const int *order, int steps, int batch, float lr,
             float *loss_log);
int nn_evaluate(void *model, void *dataset, int first, int count, float *mse);
/* Average device ms of a forward pass and of a forward + backward pass of
 * the network on an n^3 grid (device-resident synthetic input), and the
 * flops of one forward pass (2 * 27 * cin * cout per output voxel). */
int nn_benchmark(void *model, int n, int reps, double *fwd_ms, double *fwdbwd_ms);
double nn_forward_flops(int n);
/* Attach a network to a solver for POISSON_NN_FP32_DC (NULL detaches; the
 * caller keeps ownership). */
int gravity_set_nn_model(GravitySolver3D *solver, void *model);
/* Attach a start-correction network (kind 1) for POISSON_NN_FP32_DC: the
 * warm start plus its correction becomes a candidate start beside the warm
 * start and the potential guess (NULL detaches; the caller keeps ownership).
 * The grid is averaged by 2 until at most 64^3, so n must be 64 x 2^k or a
 * multiple of 8 from 16 to 64.  Rejects a network of the other kind. */
int gravity_set_nn_correction(GravitySolver3D *solver, void *model);

#ifdef __cplusplus
}
#endif
#endif /* HYDRO_MI355X_H */
