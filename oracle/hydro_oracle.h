/*
 * hydro_oracle.h -- CPU oracle for the hydro + self-gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls
 * this code: it is the checker used by tests/, by __graft_entry__.smoke() and
 * by the cpu_baseline leg of bench.py.
 *
 * It restates, in plain C11 built with -ffp-contract=off, the per-step
 * algorithm of the reference solver (hydro_solver_3d.c:162-811 and
 * gravity_solver_3d.c:164-911): ideal-gas EOS with floors, directional HLLC
 * fluxes, the CFL time step, boundary conditions, the conservative update,
 * the gravity source term, red-black SOR with a multipole boundary and the
 * central-difference acceleration.  Every function cites the reference lines
 * it follows in hydro_oracle.c.
 *
 * State is AoS, 5 doubles per cell (rho, rhou, rhov, rhow, E), flat index
 * ((i*ny)+j)*nz+k exactly as IDX3D in hydro_solver_3d.h:12.
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BC_PERIODIC = 0, OR_BC_REFLECTIVE = 1, OR_BC_OUTFLOW = 2 };
enum { OR_GRAV_PERIODIC = 0, OR_GRAV_ISOLATED = 1, OR_GRAV_DIRICHLET = 2 };
/* moment summation: 0 = sequential in cell order (the reference),
 * 1 = pairwise double-double tree over the flat cell index (the device). */
enum { OR_MOM_SEQUENTIAL = 0, OR_MOM_DDTREE = 1 };

typedef struct OrSolver OrSolver;

/* ---- pointwise pieces ---- */
void or_cons_to_prim_batch(int n, const double *U, double gamma, double *P);
void or_prim_to_cons_batch(int n, const double *P, double gamma, double *U);
void or_hllc_batch(int dir, int n, const double *UL, const double *UR,
                   double gamma, double *F);
double or_compute_dt(const double *U, int nx, int ny, int nz,
                     double dx, double dy, double dz, double cfl, double gamma);

/* ---- solver ---- */
OrSolver *or_create(int nx, int ny, int nz, double xmin, double xmax,
                    double ymin, double ymax, double zmin, double zmax,
                    double G);
void or_destroy(OrSolver *s);
void or_set_hydro(OrSolver *s, double gamma, double cfl, const int bc[6]);
void or_set_gravity(OrSolver *s, double omega, int max_iter, double tol,
                    int lmax, int moment_mode, int hoist_bc, int grav_bc);
void or_set_state(OrSolver *s, const double *U);
void or_get_state(const OrSolver *s, double *U);
void or_set_phi(OrSolver *s, const double *phi);
void or_get_phi(const OrSolver *s, double *phi);
void or_get_g(const OrSolver *s, double *gx, double *gy, double *gz);
void or_get_grav_g(const OrSolver *s, double *gx, double *gy, double *gz);
double or_get_dt(const OrSolver *s);
void or_get_diag(const OrSolver *s, int *iters, double *res, int *status);

/* One step of the reference step function (godunov_step_3d). */
int or_step(OrSolver *s);
/* The reference time loop without I/O: while (t < t_end) step; t += dt.
 * Stops early after max_steps (<=0: unlimited).  dt_log/iter_log (may be
 * NULL) receive one entry per step.  Returns the number of steps. */
int or_run(OrSolver *s, double t_end, int max_steps, double *t_out,
           double *dt_log, int *iter_log);

/* gravity_solve on an external density (potential + acceleration). */
int or_gravity_solve(OrSolver *s, const double *rho);
/* Moments as the boundary condition sees them: out[0..3] = M, xcm, ycm, zcm;
 * out[4..12] = Q[3][3] row-major.  hi-order A/B (l=3..lmax, m=0..l) go to
 * out_hi (may be NULL), laid out [l][m][A,B] with l from 3. */
void or_moments(const OrSolver *s, const double *rho, double *out,
                double *out_hi);
/* Boundary potential only: writes the six faces of phi (interior untouched). */
void or_boundary_phi(const OrSolver *s, const double *rho, double *phi);
/* Relative residual ||4 pi G rho - L_h phi||_2 / ||4 pi G rho||_2 over
 * interior cells (the tolerance measure for the defect-correction mode). */
double or_rel_residual(const OrSolver *s, const double *rho,
                       const double *phi);

/* Real regular solid harmonics C_lm, S_lm for l <= lmax (row l, m = 0..l),
 * out arrays sized (lmax+1)*(lmax+1) indexed l*(lmax+1)+m. */
void or_solid_harmonics(int lmax, double x, double y, double z, double *C,
                        double *S);

/* Pairwise double-double tree sum of n values (zero-padded to a power of 2),
 * rounded to double.  Exposed for tests. */
double or_ddtree_sum(const double *v, long n);

#ifdef __cplusplus
}
#endif
#endif
