/*
 * ref_grav_harness.c -- exposes the reference gravity solver, including its
 * static moment and boundary functions, for checking the oracle.
 *
 * TEST INFRASTRUCTURE ONLY; built into oracle/_ref/ from the reference
 * sources where they lie.
 */
#include "gravity_solver_3d.c"

static long gcells(const GravitySolver3D *s) { return (long)s->nx * s->ny * s->nz; }

void *refg_create(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                  double ymax, double zmin, double zmax, double G) {
    return gravity_solver_create(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, G);
}
void refg_destroy(void *s) { gravity_solver_destroy((GravitySolver3D *)s); }
void refg_set_params(void *s, double omega, int max_iter, double tol) {
    gravity_set_solver_params((GravitySolver3D *)s, omega, max_iter, tol);
}
void refg_set_bc(void *s, int bc) { gravity_set_boundary_type((GravitySolver3D *)s, (GravBCType)bc); }
void refg_set_phi(void *p, const double *phi) {
    GravitySolver3D *s = (GravitySolver3D *)p;
    memcpy(s->phi, phi, gcells(s) * sizeof(double));
}
int refg_solve(void *s, const double *rho) { return gravity_solve((GravitySolver3D *)s, rho); }
void refg_get(void *p, double *phi, double *gx, double *gy, double *gz) {
    GravitySolver3D *s = (GravitySolver3D *)p;
    long n = gcells(s);
    memcpy(phi, s->phi, n * sizeof(double));
    memcpy(gx, s->gx, n * sizeof(double));
    memcpy(gy, s->gy, n * sizeof(double));
    memcpy(gz, s->gz, n * sizeof(double));
}
void refg_diag(void *s, int *iters, double *res) {
    gravity_get_diagnostics((GravitySolver3D *)s, iters, res);
}
void refg_moments(void *p, const double *rho, double *out) {
    GravitySolver3D *s = (GravitySolver3D *)p;
    double Q[3][3];
    compute_mass_moments(s, rho, &out[0], &out[1], &out[2], &out[3]);
    compute_quadrupole_moment(s, rho, out[1], out[2], out[3], Q);
    memcpy(out + 4, Q, sizeof Q);
}
void refg_boundary(void *p, const double *rho, double *phi_inout) {
    GravitySolver3D *s = (GravitySolver3D *)p;
    long n = gcells(s);
    memcpy(s->phi, phi_inout, n * sizeof(double));
    s->rho = (double *)rho;
    apply_boundary_conditions(s);
    memcpy(phi_inout, s->phi, n * sizeof(double));
}
