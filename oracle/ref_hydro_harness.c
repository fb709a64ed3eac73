/*
 * ref_hydro_harness.c -- exposes the reference's static hydro step functions
 * so the oracle can be checked against the compiled reference.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
 * the reference sources where they lie (-I<reference dir>); nothing of the
 * reference is copied here.  The wrappers only move data in and out of the
 * reference's own structs and call its functions.
 */
#include "hydro_solver_3d.c"

static long ncells(const GodunovHLLC3D *s) { return (long)s->nx * s->ny * s->nz; }

void *ref_create(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                 double ymax, double zmin, double zmax, double G) {
    return solver3d_create(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, G);
}
void ref_destroy(void *s) { solver3d_destroy((GodunovHLLC3D *)s); }

void ref_set_hydro(void *p, double gamma, double cfl, const int bc[6]) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    s->gamma = gamma;
    s->cfl = cfl;
    s->bc_xmin = (BCType)bc[0]; s->bc_xmax = (BCType)bc[1];
    s->bc_ymin = (BCType)bc[2]; s->bc_ymax = (BCType)bc[3];
    s->bc_zmin = (BCType)bc[4]; s->bc_zmax = (BCType)bc[5];
}
void ref_set_grav(void *p, double omega, int max_iter, double tol) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    if (s->gravity_solver) gravity_set_solver_params(s->gravity_solver, omega, max_iter, tol);
}
void ref_set_state(void *p, const double *U) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    memcpy(s->U, U, ncells(s) * sizeof(State3D));
}
void ref_get_state(void *p, double *U) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    memcpy(U, s->U, ncells(s) * sizeof(State3D));
}
void ref_get_g(void *p, double *gx, double *gy, double *gz) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    memcpy(gx, s->gx, ncells(s) * sizeof(double));
    memcpy(gy, s->gy, ncells(s) * sizeof(double));
    memcpy(gz, s->gz, ncells(s) * sizeof(double));
}
void ref_get_phi(void *p, double *phi) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    if (s->gravity_solver) memcpy(phi, s->gravity_solver->phi, ncells(s) * sizeof(double));
}
void ref_diag(void *p, int *iters, double *res) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    *iters = s->gravity_solver ? s->gravity_solver->last_iterations : 0;
    *res = s->gravity_solver ? s->gravity_solver->last_residual : 0.0;
}
double ref_dt(void *p) { return ((GodunovHLLC3D *)p)->dt; }
void ref_step(void *p) { godunov_step_3d((GodunovHLLC3D *)p); }

/* run_simulation_3d's loop (hydro_solver_3d.c:2257-2260) without output */
int ref_run(void *p, double t_end, int max_steps, double *t_out, double *dt_log, int *iter_log) {
    GodunovHLLC3D *s = (GodunovHLLC3D *)p;
    double t = 0.0;
    int steps = 0;
    while (t < t_end) {
        godunov_step_3d(s);
        t += s->dt;
        if (dt_log) dt_log[steps] = s->dt;
        if (iter_log) iter_log[steps] = s->gravity_solver ? s->gravity_solver->last_iterations : 0;
        steps++;
        if (max_steps > 0 && steps >= max_steps) break;
    }
    if (t_out) *t_out = t;
    return steps;
}

void ref_c2p_batch(int n, const double *U, double gamma, double *P) {
    GodunovHLLC3D s; s.gamma = gamma;
    for (int c = 0; c < n; c++) {
        State3D u; memcpy(&u, U + 5 * c, sizeof u);
        Primitive3D q = cons_to_prim_3d(u, &s);
        memcpy(P + 5 * c, &q, sizeof q);
    }
}
void ref_p2c_batch(int n, const double *P, double gamma, double *U) {
    GodunovHLLC3D s; s.gamma = gamma;
    for (int c = 0; c < n; c++) {
        Primitive3D q; memcpy(&q, P + 5 * c, sizeof q);
        State3D u = prim_to_cons_3d(q, &s);
        memcpy(U + 5 * c, &u, sizeof u);
    }
}
void ref_hllc_batch(int dir, int n, const double *UL, const double *UR, double gamma, double *F) {
    GodunovHLLC3D s; s.gamma = gamma;
    for (int c = 0; c < n; c++) {
        State3D l, r, f;
        memcpy(&l, UL + 5 * c, sizeof l);
        memcpy(&r, UR + 5 * c, sizeof r);
        if (dir == 0) f = hllc_riemann_solver_x_3d(l, r, &s);
        else if (dir == 1) f = hllc_riemann_solver_y_3d(l, r, &s);
        else f = hllc_riemann_solver_z_3d(l, r, &s);
        memcpy(F + 5 * c, &f, sizeof f);
    }
}
double ref_compute_dt(const double *U, int nx, int ny, int nz, double dx, double dy, double dz,
                      double cfl, double gamma) {
    GodunovHLLC3D s;
    s.nx = nx; s.ny = ny; s.nz = nz; s.dx = dx; s.dy = dy; s.dz = dz;
    s.cfl = cfl; s.gamma = gamma; s.U = (State3D *)U;
    compute_dt_3d(&s);
    return s.dt;
}

/* The reference's initial-condition problems on its own solver (test
 * infrastructure: compared with hydro_fill_initial_state). */
int ref_ic_problem(void *p, int problem) {
    set_initial_conditions_3d((GodunovHLLC3D *)p, (ProblemType)problem);
    return 0;
}
int ref_ic_deck(void *p, const char *deck_path) {
    InputDeck *d = input_deck_create(deck_path);
    if (!d) return -1;
    set_initial_conditions_3d_from_deck((GodunovHLLC3D *)p, d);
    input_deck_destroy(d);
    return 0;
}
void ref_set_gamma(void *p, double gamma) { ((GodunovHLLC3D *)p)->gamma = gamma; }
/* main_3d.c's set-up without output: the solver and its initial state from a
 * deck file (hydro_solver_3d.c:130-156, 2092-2133). */
void *ref_from_deck(const char *deck_path) {
    InputDeck *d = input_deck_create(deck_path);
    if (!d) return NULL;
    GodunovHLLC3D *s = solver3d_create_from_deck(d);
    if (s) set_initial_conditions_3d_from_deck(s, d);
    input_deck_destroy(d);
    return s;
}
