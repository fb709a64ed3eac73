/*
 * hydro_solver_3d.h -- first-order HLLC Godunov solver for 3D Euler gas
 * dynamics with optional self-gravity, running on one MI355X.
 *
 * Drop-in for the reference's hydro_solver_3d.h (lines 1-97): the macros,
 * State3D, Primitive3D and the ten functions keep their definitions and
 * signatures; GodunovHLLC3D keeps every field in the reference's order and
 * appends a device context.  State lives on the GPU between steps; the host
 * array U is the caller-visible mirror, uploaded when a run starts and
 * downloaded at outputs and when a run ends.
 *
 * U_half, F_x, F_y, F_z stay in the struct for layout compatibility but are
 * NULL: the fluxes never leave the GPU's on-chip memory.
 */
#ifndef HYDRO_SOLVER_3D_H
#define HYDRO_SOLVER_3D_H

#include <stdbool.h>
#include <stdio.h>

#include "gravity_solver_3d.h"
#include "input_deck.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hydro_solver_3d.h:10-12 */
#define RHO_FLOOR 1.0e-5
#define P_FLOOR 1.0e-6
#define IDX3D(i, j, k, ny, nz) (((i) * (ny) + (j)) * (nz) + (k))

/* hydro_solver_3d.h:15-21 -- 40 bytes, AoS on the host */
typedef struct {
    double rho;
    double rhou;
    double rhov;
    double rhow;
    double E;
} State3D;

/* hydro_solver_3d.h:24-30 */
typedef struct {
    double rho;
    double u;
    double v;
    double w;
    double p;
} Primitive3D;

struct HydroDevice; /* opaque: SoA state planes, boundary buffers, dt control */

/* hydro_solver_3d.h:33-63, then the fields this build adds. */
typedef struct GodunovHLLC3D {
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double dx, dy, dz;
    double dt;
    double cfl;

    BCType bc_xmin, bc_xmax;
    BCType bc_ymin, bc_ymax;
    BCType bc_zmin, bc_zmax;

    ProblemType problem;
    double gamma;

    State3D *U;
    State3D *U_half; /* NULL */
    State3D *F_x;    /* NULL */
    State3D *F_y;    /* NULL */
    State3D *F_z;    /* NULL */

    double *gx;
    double *gy;
    double *gz;

    GravitySolver3D *gravity_solver;
    bool enable_gravity;
    double G_const;

    /* ---- added by this build ---- */
    struct HydroDevice *dev;
    int max_steps;        /* run loops stop after this many steps (0: none) */
    bool output_enable;   /* run loops write CSV/VTK at output times        */
    long steps_taken;     /* steps since the state was last uploaded        */
    int ic_status;        /* 0, or -1 when the deck's ic_file failed to load:
                             U is left as it was and the run loops refuse
                             to start from it                               */
    bool checkpoint_enable; /* run loops also write output/checkpoint_N.bin
                               (HYDCK001) at output times                  */
} GodunovHLLC3D;

/* hydro_solver_3d.h:69-73 */
GodunovHLLC3D *solver3d_create(int nx, int ny, int nz, double xmin,
                               double xmax, double ymin, double ymax,
                               double zmin, double zmax, double G);
/* hydro_solver_3d.h:74 */
void solver3d_destroy(GodunovHLLC3D *solver);
/* hydro_solver_3d.h:77-78 */
void set_initial_conditions_3d(GodunovHLLC3D *solver, ProblemType problem);
void set_initial_conditions_3d_from_deck(GodunovHLLC3D *solver,
                                         const InputDeck *deck);
/* hydro_solver_3d.h:81 */
void run_simulation_3d(GodunovHLLC3D *solver, double t_end, double output_dt);
/* hydro_solver_3d.h:84-85 */
void run_simulation_3d_restart(GodunovHLLC3D *solver, double t_start,
                               double t_end, double output_dt,
                               int start_output_count);
/* hydro_solver_3d.h:88-89 */
void write_output_to_file_3d(const GodunovHLLC3D *solver, int step);
void write_output_vtk(const GodunovHLLC3D *solver, int step);
/* hydro_solver_3d.h:92 */
int load_from_csv_restart(GodunovHLLC3D *solver, const char *filename);
/* hydro_solver_3d.h:95 */
GodunovHLLC3D *solver3d_create_from_deck(const InputDeck *deck);

#ifdef __cplusplus
}
#endif
#endif /* HYDRO_SOLVER_3D_H */
