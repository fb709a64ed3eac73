/*
 * gravity_solver_3d.h -- self-gravity: Poisson solve and acceleration.
 *
 * Drop-in for the reference's gravity_solver_3d.h (lines 1-151).  GravBCType
 * and the nine functions keep their signatures; GravitySolver3D keeps every
 * field in the reference's order (callers read phi/gx/gy/gz and set rho
 * directly, hydro_solver_3d.c:585-587) and appends the device context and
 * the settings this build adds.  The solve runs on the GPU; the host arrays
 * phi/gx/gy/gz are refreshed when a function below returns.
 */
#ifndef GRAVITY_SOLVER_3D_H
#define GRAVITY_SOLVER_3D_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gravity_solver_3d.h:7 */
#define IDX3D_GRAV(i, j, k, ny, nz) (((i) * (ny) + (j)) * (nz) + (k))

/* gravity_solver_3d.h:10-14 */
typedef enum {
    GRAV_BC_PERIODIC,
    GRAV_BC_ISOLATED,
    GRAV_BC_DIRICHLET
} GravBCType;

struct GravDevice; /* opaque: device arrays, streams, solve control */

/* gravity_solver_3d.h:17-43, then the fields this build adds. */
typedef struct GravitySolver3D {
    int nx, ny, nz;
    double dx, dy, dz;
    double xmin, xmax;
    double ymin, ymax;
    double zmin, zmax;

    double G;

    double *phi;
    double *rho;
    double *gx;
    double *gy;
    double *gz;

    double omega;
    int max_iterations;
    double tolerance;

    GravBCType bc_type;

    int last_iterations;
    double last_residual;

    /* ---- added by this build ---- */
    struct GravDevice *dev;
    int multipole_lmax;  /* 2 = monopole + quadrupole (the reference); <= 8  */
    int poisson_mode;    /* PoissonMode from input_deck.h                    */
    int criterion;       /* SorCriterion from input_deck.h                   */
    int last_status;     /* 0 converged, -1 not converged, < -1 device error */
    double last_rel_residual; /* ||f - L phi|| / ||f|| after the last solve  */
    double last_solve_ms;     /* device time of the last potential solve     */
    void *nn_model;      /* learned-guess network (nn_model_create), or NULL */
    int last_guess;      /* POISSON_NN_FP32_DC: 1 network guess used, 0 warm
                            start kept, -1 no guess possible (2 below)    */
    double last_guess_rel_residual; /* residual of the network guess         */
    double last_warm_rel_residual;  /* residual of the warm start            */
    int dc_inner;        /* DcInner from input_deck.h: inner solver of the FP32
                            modes; with DC_INNER_MULTIGRID, last_iterations
                            counts V-cycles                               */
    int moment_order;    /* MomentOrder from input_deck.h                    */
    void *nn_correction; /* start-correction network (nn_model_set_kind 1),
                            or NULL; with it last_guess 2 = warm start plus
                            the network's correction used                 */
    double last_corr_rel_residual; /* residual of the corrected warm start */
} GravitySolver3D;

/* gravity_solver_3d.h:64-68 */
GravitySolver3D *gravity_solver_create(int nx, int ny, int nz, double xmin,
                                       double xmax, double ymin, double ymax,
                                       double zmin, double zmax, double G);
/* gravity_solver_3d.h:75 */
void gravity_solver_destroy(GravitySolver3D *solver);
/* gravity_solver_3d.h:87; rho_field is a host array of nx*ny*nz doubles */
int gravity_solve_potential(GravitySolver3D *solver, const double *rho_field);
/* gravity_solver_3d.h:97 */
void gravity_compute_acceleration(GravitySolver3D *solver);
/* gravity_solver_3d.h:107-110 */
void gravity_set_solver_params(GravitySolver3D *solver, double omega,
                               int max_iter, double tolerance);
/* gravity_solver_3d.h:118 */
void gravity_set_boundary_type(GravitySolver3D *solver, GravBCType bc_type);
/* gravity_solver_3d.h:127-129 */
void gravity_get_diagnostics(const GravitySolver3D *solver, int *iterations,
                             double *residual);
/* gravity_solver_3d.h:141 */
int gravity_solve(GravitySolver3D *solver, const double *rho_field);
/* gravity_solver_3d.h:149 */
void gravity_write_potential(const GravitySolver3D *solver,
                             const char *filename);

#ifdef __cplusplus
}
#endif
#endif /* GRAVITY_SOLVER_3D_H */
