// hydro_api.hip -- the hydro_solver_3d.h API on the GPU.
//
// Lifecycle follows hydro_solver_3d.c:33-156, the step godunov_step_3d
// (:637-811) runs as two kernels per step (hydro_kernels.hip) after the
// device gravity solve, and the time loops follow :2241-2330.  State stays
// on the device between steps; solver->U is the host mirror.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>

#include "../../include/hydro_mi355x.h"
#include "host_common.hpp"

// One non-blocking stream (default priority) for all work of the library.
hipStream_t hyd_stream() {
    static hipStream_t st = nullptr;
    static bool tried = false;
    if (!tried) {
        tried = true;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            fprintf(stderr, "hydro_mi355x: no usable GPU (%s)\n", hipGetErrorString(e));
            st = nullptr;
        }
    }
    return st;
}

namespace {

void dev_free(HydroDevice* d) {
    if (!d) return;
    hyd_output_free(d);
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f)
            if (d->U[b][f]) (void)hipFree(d->U[b][f]);
    void* ptrs[] = {d->face, d->ms, d->dt_log, d->g_zero, d->aos};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (d->ev0) (void)hipEventDestroy(d->ev0);
    if (d->ev1) (void)hipEventDestroy(d->ev1);
    delete d;
}

int dev_create(GodunovHLLC3D* s) {
    HydroDevice* d = new HydroDevice();
    s->dev = d;
    d->nx = s->nx; d->ny = s->ny; d->nz = s->nz;
    d->n = (long)s->nx * s->ny * s->nz;
    d->stream = hyd_stream();
    if (!d->stream) return -2;
    // x planes per step-kernel block; unset or 0 lets the launcher choose
    if (const char* e = getenv("HYDRO_MI355X_CX")) d->cx = atoi(e) > 0 ? atoi(e) : 0;
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f) HYD_CHECK(hipMalloc((void**)&d->U[b][f], sizeof(double) * d->n));
    HYD_CHECK(hipMalloc((void**)&d->face, sizeof(double) * hyd::face_buffer_len(d->nx, d->ny, d->nz)));
    HYD_CHECK(hipMalloc((void**)&d->ms, sizeof(unsigned long long) * HydroDevice::MS_WORDS));
    HYD_CHECK(hipMalloc((void**)&d->dt_log, sizeof(double) * HydroDevice::DT_LOG));
    HYD_CHECK(hipMalloc((void**)&d->aos, sizeof(double) * 5 * d->n));
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f)
            HYD_CHECK(hipMemsetAsync(d->U[b][f], 0, sizeof(double) * d->n, d->stream));
    HYD_CHECK(hipMemsetAsync(d->ms, 0, sizeof(unsigned long long) * HydroDevice::MS_WORDS,
                             d->stream));
    HYD_CHECK(hipEventCreate(&d->ev0));
    HYD_CHECK(hipEventCreate(&d->ev1));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int upload(GodunovHLLC3D* s) {
    HydroDevice* d = s->dev;
    HYD_CHECK(hipMemcpyAsync(d->aos, s->U, sizeof(State3D) * d->n, hipMemcpyHostToDevice, d->stream));
    HYD_CHECK(hyd::launch_aos_to_soa(d->aos, d->U[d->cur], d->n, d->stream));
    HYD_CHECK(hipMemsetAsync(d->ms, 0, sizeof(unsigned long long) * HydroDevice::MS_WORDS,
                             d->stream));
    HYD_CHECK(hyd::launch_hydro_maxspeed(d->U[d->cur], d->n, s->gamma, d->ms, d->ms + 3,
                                         d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    d->step = 0;
    d->uploaded = true;
    s->steps_taken = 0;
    return 0;
}

// Device state to the host mirror.  Until the mirror has been uploaded (after
// set_initial_conditions_3d, load_from_csv_restart, a checkpoint read) the
// mirror is the state, and this leaves it alone.
int download(GodunovHLLC3D* s) {
    HydroDevice* d = s->dev;
    if (!d->uploaded) return 0;
    HYD_CHECK(hyd::launch_soa_to_aos(d->aos, d->U[d->cur], d->n, d->stream));
    HYD_CHECK(hipMemcpyAsync(s->U, d->aos, sizeof(State3D) * d->n, hipMemcpyDeviceToHost, d->stream));
    if (s->enable_gravity && s->gravity_solver) {
        // hydro_solver_3d.c:584-588: the hydro copies of g
        const size_t b = sizeof(double) * d->n;
        HYD_CHECK(hipMemcpyAsync(s->gx, grav_internal_g(s->gravity_solver, 0), b,
                                 hipMemcpyDeviceToHost, d->stream));
        HYD_CHECK(hipMemcpyAsync(s->gy, grav_internal_g(s->gravity_solver, 1), b,
                                 hipMemcpyDeviceToHost, d->stream));
        HYD_CHECK(hipMemcpyAsync(s->gz, grav_internal_g(s->gravity_solver, 2), b,
                                 hipMemcpyDeviceToHost, d->stream));
    }
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

// godunov_step_3d on the device; no host synchronisation unless the gravity
// solve needs one.  Returns the gravity status (0, -1) or < -1 on error.
int step_device(GodunovHLLC3D* s) {
    HydroDevice* d = s->dev;
    int gst = 0;
    const bool grav = s->enable_gravity && s->gravity_solver;
    if (grav && !d->skip_gravity) {
        // solve_gravity_field, hydro_solver_3d.c:558-591: the density is the
        // raw conserved rho, here the rho plane itself
        gst = grav_internal_solve(s->gravity_solver, d->U[d->cur][0], d->stream);
        if (gst < -1) return gst;
        if (gst != 0) fprintf(stderr, "Warning: Gravity solver did not converge fully.\n");
    }
    hyd::HydroStepArgs a;
    for (int f = 0; f < 5; ++f) {
        a.uin[f] = d->U[d->cur][f];
        a.uout[f] = d->U[1 - d->cur][f];
    }
    for (int c = 0; c < 3; ++c) a.g[c] = grav ? grav_internal_g(s->gravity_solver, c) : nullptr;
    a.face = d->face;
    a.ms_in = d->ms + (d->step % 3);
    a.ms_out = d->ms + ((d->step + 1) % 3);
    a.ms_clear = d->ms + ((d->step + 2) % 3);
    a.nan_in = a.ms_in + 3;
    a.nan_out = a.ms_out + 3;
    a.nan_clear = a.ms_clear + 3;
    a.dt_out = d->dt_log + (d->step % HydroDevice::DT_LOG);
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz;
    a.cx = d->cx;
    a.ntj = a.ntk = a.wpb = 0;
    a.dx = s->dx; a.dy = s->dy; a.dz = s->dz;
    a.cfl = s->cfl;
    a.gamma = s->gamma;
    a.gravity = grav ? 1 : 0;
    a.zcopy = hyd::hydro_zcopy(s->nz, s->bc_zmin, s->bc_zmax) ? 1 : 0;
    HYD_CHECK(hipEventRecord(d->record_ev[0] ? d->record_ev[0] : d->ev0, d->stream));
    HYD_CHECK(hyd::launch_hydro_step(a, d->stream));
    if (d->record_ev[1]) HYD_CHECK(hipEventRecord(d->record_ev[1], d->stream));
    hyd::HydroBCArgs bc;
    for (int f = 0; f < 5; ++f) bc.u[f] = d->U[1 - d->cur][f];
    bc.face = d->face;
    bc.ms_out = a.ms_out;
    bc.nan_out = a.nan_out;
    bc.nx = s->nx; bc.ny = s->ny; bc.nz = s->nz;
    bc.bc[0] = s->bc_xmin; bc.bc[1] = s->bc_xmax;
    bc.bc[2] = s->bc_ymin; bc.bc[3] = s->bc_ymax;
    bc.bc[4] = s->bc_zmin; bc.bc[5] = s->bc_zmax;
    bc.gamma = s->gamma;
    bc.zcopy = a.zcopy;
    HYD_CHECK(hyd::launch_hydro_bc(bc, d->stream));
    HYD_CHECK(hipEventRecord(d->ev1, d->stream));
    if (d->record_ev[2]) HYD_CHECK(hipEventRecord(d->record_ev[2], d->stream));
    d->cur = 1 - d->cur;
    d->step++;
    s->steps_taken++;
    return gst;
}

int read_dt(GodunovHLLC3D* s, long step, double* dt) {
    HydroDevice* d = s->dev;
    HYD_CHECK(hipMemcpyAsync(dt, d->dt_log + (step % HydroDevice::DT_LOG), sizeof(double),
                             hipMemcpyDeviceToHost, d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

// dt of `count` consecutive steps from `first` (count <= DT_LOG): one or two
// contiguous pieces of the ring.
int read_dt_range(GodunovHLLC3D* s, long first, int count, double* out) {
    HydroDevice* d = s->dev;
    const long head = first % HydroDevice::DT_LOG;
    const long a = std::min<long>(count, HydroDevice::DT_LOG - head);
    HYD_CHECK(hipMemcpyAsync(out, d->dt_log + head, sizeof(double) * a, hipMemcpyDeviceToHost,
                             d->stream));
    if (count > a)
        HYD_CHECK(hipMemcpyAsync(out + a, d->dt_log, sizeof(double) * (count - a),
                                 hipMemcpyDeviceToHost, d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int ensure_uploaded(GodunovHLLC3D* s) {
    if (!s->dev->uploaded) return upload(s);
    return 0;
}

void mkdir_output() { mkdir("output", 0755); }

void write_checkpoint_file(GodunovHLLC3D* s, double t, int count) {
    char name[256];
    snprintf(name, sizeof name, "output/checkpoint_%06d.bin", count);
    (void)solver3d_write_checkpoint(s, name, t, count);
}

// run_simulation_3d / run_simulation_3d_restart, hydro_solver_3d.c:2241-2330.
// The output at the start comes from the host mirror (it may hold g read
// from a restart file); outputs after steps are device snapshots written
// while the steps go on; the final one follows the last download.
void run_loop(GodunovHLLC3D* s, double t_start, double t_end, double output_dt, int output_count,
              double next_output_time) {
    if (s->ic_status != 0) {
        fprintf(stderr,
                "Error: the initial state was not loaded (ic_file); the run does not start.\n");
        return;
    }
    double t = t_start;
    int step_counter = 0;
    if (upload(s)) return;
    if (s->output_enable) {
        mkdir_output();
        write_output_to_file_3d(s, output_count);
        write_output_vtk(s, output_count);
        if (s->checkpoint_enable) write_checkpoint_file(s, t, output_count);
    }
    output_count++;
    while (t < t_end) {
        if (step_device(s) < -1) break;
        if (read_dt(s, s->dev->step - 1, &s->dt)) break;
        t += s->dt;
        step_counter++;
        if (t >= next_output_time) {
            if (s->output_enable) {
                if (hyd_output_async(s, output_count)) break;
                if (s->checkpoint_enable) write_checkpoint_file(s, t, output_count);
            }
            printf("Step %d, t = %.5f, dt = %.5e, Output %d\n", step_counter, t, s->dt, output_count);
            output_count++;
            next_output_time += output_dt;
        }
        if (step_counter % 10 == 0) printf("Step %d, t = %.5f, dt = %.5e\n", step_counter, t, s->dt);
        if (s->max_steps > 0 && step_counter >= s->max_steps) break;
    }
    hyd_output_drain(s->dev);
    if (download(s)) return;
    if (s->output_enable && t - (next_output_time - output_dt) > 1e-6) {
        write_output_to_file_3d(s, output_count);
        write_output_vtk(s, output_count);
        if (s->checkpoint_enable) write_checkpoint_file(s, t, output_count);
    }
    printf("Simulation completed: %d steps, final time: %.5f\n", step_counter, t);
}

}  // namespace

int hyd_download(GodunovHLLC3D* s) { return download(s); }

extern "C" {

int hydro_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

GodunovHLLC3D* solver3d_create(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                               double ymax, double zmin, double zmax, double G) {
    if (nx < 3 || ny < 3 || nz < 3) {
        fprintf(stderr, "Error: the device solver needs nx, ny, nz >= 3 (got %d x %d x %d).\n", nx,
                ny, nz);
        return NULL;
    }
    GodunovHLLC3D* s = (GodunovHLLC3D*)calloc(1, sizeof(GodunovHLLC3D));
    if (!s) {
        fprintf(stderr, "Failed to allocate solver structure.\n");
        return NULL;
    }
    // hydro_solver_3d.c:44-63
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->xmin = xmin; s->xmax = xmax; s->ymin = ymin; s->ymax = ymax; s->zmin = zmin; s->zmax = zmax;
    s->dx = (xmax - xmin) / nx;
    s->dy = (ymax - ymin) / ny;
    s->dz = (zmax - zmin) / nz;
    s->cfl = 0.4;
    s->dt = 0.0;
    s->gamma = 5.0 / 3.0;
    s->G_const = G;
    s->enable_gravity = (G > 0.0);
    s->output_enable = true;
    s->gravity_solver = NULL;
    if (s->enable_gravity) {  // hydro_solver_3d.c:65-82
        s->gravity_solver = gravity_solver_create(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, G);
        if (!s->gravity_solver) {
            fprintf(stderr, "Warning: Failed to create gravity solver. Disabling gravity.\n");
            s->enable_gravity = false;
        } else {
            gravity_set_boundary_type(s->gravity_solver, GRAV_BC_ISOLATED);
            printf("Gravity solver initialized with G = %.6e\n", G);
        }
    }
    const long n = (long)nx * ny * nz;
    s->U = (State3D*)calloc(n, sizeof(State3D));
    s->gx = (double*)calloc(n, sizeof(double));
    s->gy = (double*)calloc(n, sizeof(double));
    s->gz = (double*)calloc(n, sizeof(double));
    if (!s->U || !s->gx || !s->gy || !s->gz || dev_create(s) != 0) {
        fprintf(stderr, "Failed to allocate solver arrays.\n");
        solver3d_destroy(s);
        return NULL;
    }
    return s;
}

void solver3d_destroy(GodunovHLLC3D* s) {
    if (!s) return;
    free(s->U);
    free(s->gx);
    free(s->gy);
    free(s->gz);
    if (s->gravity_solver) gravity_solver_destroy(s->gravity_solver);
    dev_free(s->dev);
    free(s);
}

GodunovHLLC3D* solver3d_create_from_deck(const InputDeck* deck) {
    // hydro_solver_3d.c:130-156, plus the keys this build adds
    GodunovHLLC3D* s = solver3d_create(deck->nx, deck->ny, deck->nz, deck->xmin, deck->xmax,
                                       deck->ymin, deck->ymax, deck->zmin, deck->zmax,
                                       deck->G_const);
    if (!s) return NULL;
    s->cfl = deck->cfl;
    s->gamma = deck->gamma;
    s->problem = deck->problem;
    s->bc_xmin = deck->bc_xmin; s->bc_xmax = deck->bc_xmax;
    s->bc_ymin = deck->bc_ymin; s->bc_ymax = deck->bc_ymax;
    s->bc_zmin = deck->bc_zmin; s->bc_zmax = deck->bc_zmax;
    s->max_steps = deck->max_steps;
    s->output_enable = deck->output_enable;
    s->checkpoint_enable = deck->checkpoint;
    if (s->gravity_solver) {
        gravity_set_solver_params(s->gravity_solver, deck->sor_omega, deck->sor_max_iter,
                                  deck->sor_tol);
        gravity_set_multipole_lmax(s->gravity_solver, deck->multipole_lmax);
        gravity_set_poisson_mode(s->gravity_solver, deck->poisson_mode);
        if (deck->dc_inner != DC_INNER_SOR &&
            gravity_set_dc_inner(s->gravity_solver, deck->dc_inner) != 0)
            fprintf(stderr, "Warning: dc_inner ignored; the FP32 modes use SOR sweeps.\n");
        gravity_set_criterion(s->gravity_solver, deck->sor_criterion);
        gravity_set_moment_order(s->gravity_solver, deck->moment_order);
        if (deck->nn_guess[0]) {
            void* m = nn_model_create(0);
            if (m && nn_model_load(m, deck->nn_guess) == 0) {
                gravity_set_nn_model(s->gravity_solver, m);
                s->gravity_solver->dev->owns_nn = true;
            } else {
                fprintf(stderr,
                        "Warning: cannot load the guess network %s; the Poisson solve starts "
                        "from the previous potential.\n",
                        deck->nn_guess);
                nn_model_destroy(m);
            }
        }
        if (deck->nn_correction[0]) {
            void* m = nn_model_create(0);
            if (m && nn_model_load(m, deck->nn_correction) == 0 &&
                gravity_set_nn_correction(s->gravity_solver, m) == 0) {
                s->gravity_solver->dev->owns_corr = true;
            } else {
                fprintf(stderr,
                        "Warning: cannot load the start-correction network %s; the Poisson "
                        "solve does without it.\n",
                        deck->nn_correction);
                nn_model_destroy(m);
            }
        }
    }
    return s;
}

// The reference's problem generators run on the host (problems.cpp); the
// state reaches the device at the next upload.
void set_initial_conditions_3d(GodunovHLLC3D* s, ProblemType problem) {
    s->problem = problem;
    s->dev->uploaded = false;
    s->ic_status = 0;
    if (hydro_fill_initial_state(s->nx, s->ny, s->nz, s->xmin, s->xmax, s->ymin, s->ymax, s->zmin,
                                 s->zmax, s->gamma, s->G_const, (int)problem, NULL, s->U) != 0)
        fprintf(stderr,
                "set_initial_conditions_3d: problem type %d has no generator (as in the "
                "reference); the state is left as it was.\n",
                (int)problem);
}

// With an ic_file the state comes from that file alone: when it cannot be
// loaded U is left as it was, ic_status is set to -1 and the run loops refuse
// to start (the problem generator does not stand in for the file).
void set_initial_conditions_3d_from_deck(GodunovHLLC3D* s, const InputDeck* deck) {
    s->problem = deck->problem;
    if (deck->ic_file[0]) {
        s->ic_status = hyd_read_state_file(s, deck->ic_file, nullptr, nullptr) == 0 ? 0 : -1;
        return;
    }
    s->dev->uploaded = false;
    s->ic_status = 0;
    if (hydro_fill_initial_state(s->nx, s->ny, s->nz, s->xmin, s->xmax, s->ymin, s->ymax, s->zmin,
                                 s->zmax, s->gamma, s->G_const, (int)deck->problem, deck,
                                 s->U) != 0)
        fprintf(stderr,
                "set_initial_conditions_3d_from_deck: problem type %d has no generator (as in "
                "the reference); the state is left as it was.\n",
                (int)deck->problem);
}

void run_simulation_3d(GodunovHLLC3D* s, double t_end, double output_dt) {
    printf("Starting 3D simulation from t=0.0 to t=%.5f\n", t_end);
    printf("Grid: %d x %d x %d = %d cells\n", s->nx, s->ny, s->nz, s->nx * s->ny * s->nz);
    run_loop(s, 0.0, t_end, output_dt, 0, output_dt);
}

void run_simulation_3d_restart(GodunovHLLC3D* s, double t_start, double t_end, double output_dt,
                               int start_output_count) {
    printf("Restarting 3D simulation from t=%.5f to t=%.5f\n", t_start, t_end);
    printf("Starting output count: %d\n", start_output_count);
    printf("Grid: %d x %d x %d = %d cells\n", s->nx, s->ny, s->nz, s->nx * s->ny * s->nz);
    run_loop(s, t_start, t_end, output_dt, start_output_count, t_start + output_dt);
}

// ------------------------------------------------------------- extensions --
int solver3d_upload(GodunovHLLC3D* s) { return upload(s); }
int solver3d_download(GodunovHLLC3D* s) { return download(s); }

int solver3d_set_state(GodunovHLLC3D* s, const double* U_aos) {
    memcpy(s->U, U_aos, sizeof(State3D) * (size_t)s->nx * s->ny * s->nz);
    s->ic_status = 0;
    return upload(s);
}

int solver3d_step(GodunovHLLC3D* s) {
    if (ensure_uploaded(s)) return -2;
    int st = step_device(s);
    if (st < -1) return st;
    if (read_dt(s, s->dev->step - 1, &s->dt)) return -2;
    return 0;
}

// The device keeps the last DT_LOG steps' dt in a ring: steps go in rounds
// of at most DT_LOG, each round's dt read back before the next round starts.
int solver3d_run_steps(GodunovHLLC3D* s, int nsteps, double* dt_log) {
    if (ensure_uploaded(s)) return -2;
    HydroDevice* d = s->dev;
    for (int done = 0; done < nsteps;) {
        const int round = std::min(nsteps - done, (int)HydroDevice::DT_LOG);
        const long first = d->step;
        for (int q = 0; q < round; ++q)
            if (step_device(s) < -1) return -2;
        if (dt_log && read_dt_range(s, first, round, dt_log + done)) return -2;
        done += round;
    }
    if (nsteps > 0 && read_dt(s, d->step - 1, &s->dt)) return -2;
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int solver3d_run(GodunovHLLC3D* s, double t_end, int max_steps, double* t_out, double* dt_log,
                 int* iter_log, int log_cap) {
    if (ensure_uploaded(s)) return -2;
    double t = 0.0;
    int steps = 0;
    while (t < t_end) {
        if (step_device(s) < -1) return -2;
        if (read_dt(s, s->dev->step - 1, &s->dt)) return -2;
        t += s->dt;
        if (steps < log_cap) {
            if (dt_log) dt_log[steps] = s->dt;
            if (iter_log)
                iter_log[steps] = (s->enable_gravity && s->gravity_solver)
                                      ? s->gravity_solver->last_iterations
                                      : 0;
        }
        steps++;
        if (max_steps > 0 && steps >= max_steps) break;
    }
    if (t_out) *t_out = t;
    return steps;
}

int solver3d_benchmark_steps(GodunovHLLC3D* s, int nsteps, double* wall_ms, double* step_ms,
                             double* bc_ms) {
    if (ensure_uploaded(s)) return -2;
    HydroDevice* d = s->dev;
    if (nsteps <= 0) return -1;
    hipEvent_t* ev = (hipEvent_t*)calloc(3 * (size_t)nsteps + 1, sizeof(hipEvent_t));
    if (!ev) return -2;
    int rc = 0;
    for (int q = 0; q < 3 * nsteps + 1 && rc == 0; ++q)
        if (hipEventCreate(&ev[q]) != hipSuccess) rc = -2;
    const bool grav = s->enable_gravity && s->gravity_solver;
    if (rc == 0 && hipStreamSynchronize(d->stream) != hipSuccess) rc = -2;
    if (rc == 0) (void)hipEventRecord(ev[0], d->stream);
    for (int q = 0; q < nsteps && rc == 0; ++q) {
        // the step kernel and the boundary kernel bracketed by their own events
        if (grav) {
            int gst = grav_internal_solve(s->gravity_solver, d->U[d->cur][0], d->stream);
            if (gst < -1) rc = gst;
        }
        d->record_ev[0] = ev[3 * q + 1];
        d->record_ev[1] = ev[3 * q + 2];
        d->record_ev[2] = ev[3 * q + 3];
        d->skip_gravity = true;
        if (rc == 0 && step_device(s) < -1) rc = -2;
        d->skip_gravity = false;
        d->record_ev[0] = d->record_ev[1] = d->record_ev[2] = nullptr;
    }
    if (rc == 0 && hipStreamSynchronize(d->stream) != hipSuccess) rc = -2;
    if (rc == 0) {
        float ms = 0.0f;
        double st = 0.0, bt = 0.0;
        (void)hipEventElapsedTime(&ms, ev[0], ev[3 * nsteps]);
        if (wall_ms) *wall_ms = ms;
        for (int q = 0; q < nsteps; ++q) {
            (void)hipEventElapsedTime(&ms, ev[3 * q + 1], ev[3 * q + 2]);
            st += ms;
            (void)hipEventElapsedTime(&ms, ev[3 * q + 2], ev[3 * q + 3]);
            bt += ms;
        }
        if (step_ms) *step_ms = st / nsteps;
        if (bc_ms) *bc_ms = bt / nsteps;
        (void)read_dt(s, d->step - 1, &s->dt);
    }
    for (int q = 0; q < 3 * nsteps + 1; ++q)
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    free(ev);
    return rc;
}

double hydro_sequential_sum(const double* x, long n, int method, long long* stats, int* status) {
    hipStream_t st = hyd_stream();
    *status = -2;
    if (!st || n < 0) return 0.0;
    const long npad = hyd::seqsum_npad(n > 0 ? n : 1);
    const long len = npad + hyd::seqsum_workspace_len(npad, 1) + 2 + 9;
    double* buf = nullptr;
    double out = 0.0;
    long long* dstats = nullptr;
    if (hipMalloc((void**)&buf, sizeof(double) * len) == hipSuccess) {
        dstats = stats ? reinterpret_cast<long long*>(buf + len - 9) : nullptr;
        if (hipMemsetAsync(buf, 0, sizeof(double) * npad, st) == hipSuccess &&
            (n == 0 || hipMemcpyAsync(buf, x, sizeof(double) * n, hipMemcpyHostToDevice, st) ==
                           hipSuccess) &&
            hyd::launch_seqsum(method, buf, npad, 1, buf + npad, buf + len - 11, st, dstats) ==
                hipSuccess &&
            hipMemcpyAsync(&out, buf + len - 11, sizeof out, hipMemcpyDeviceToHost, st) ==
                hipSuccess &&
            (!stats || method == hyd::SEQSUM_CHAIN ||
             hipMemcpyAsync(stats, dstats, 9 * sizeof(long long), hipMemcpyDeviceToHost, st) ==
                 hipSuccess) &&
            hipStreamSynchronize(st) == hipSuccess)
            *status = 0;
    }
    if (buf) (void)hipFree(buf);
    return out;
}

int solver3d_last_timing(const GodunovHLLC3D* s, double* hydro_ms, double* gravity_ms) {
    HydroDevice* d = s->dev;
    float ms = 0.0f;
    HYD_CHECK(hipEventSynchronize(d->ev1));
    HYD_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    if (hydro_ms) *hydro_ms = ms;
    if (gravity_ms) *gravity_ms = s->gravity_solver ? s->gravity_solver->last_solve_ms : 0.0;
    return 0;
}

}  // extern "C"
