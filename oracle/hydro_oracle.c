/*
 * hydro_oracle.c -- CPU oracle for the hydro + self-gravity step.
 *
 * TEST INFRASTRUCTURE ONLY (see hydro_oracle.h).  Build with
 * -O2 -std=c11 -ffp-contract=off: the reference is compiled by gcc under
 * -std=c11, which emits no fused multiply-adds, so every expression below is
 * evaluated as written, left to right, one rounding per operation.
 *
 * Each function names the reference lines whose behaviour it restates.  The
 * arithmetic order inside every expression is part of the contract with the
 * reference and is kept as the reference evaluates it.
 */
#include "hydro_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_RHO_FLOOR 1.0e-5 /* hydro_solver_3d.h:10 */
#define OR_P_FLOOR 1.0e-6   /* hydro_solver_3d.h:11 */
#define OR_PI 3.14159265358979323846
#define OR_LMAX_MAX 8

/* hydro_solver_3d.c:11-12 -- ternaries, a NaN first argument yields b. */
static inline double omin(double a, double b) { return (a < b) ? a : b; }
static inline double omax(double a, double b) { return (a > b) ? a : b; }

typedef struct { double d, mx, my, mz, e; } Cons;
typedef struct { double d, u, v, w, p; } Prim;

struct OrSolver {
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double dx, dy, dz, dt, cfl, gamma, G;
    int bc[6]; /* xmin xmax ymin ymax zmin zmax */
    int gravity;
    Cons *U, *Fx, *Fy, *Fz;
    double *hgx, *hgy, *hgz; /* hydro-side copies of g */
    /* gravity solver state (gravity_solver_3d.c:49-103) */
    double *phi, *gx, *gy, *gz;
    double omega, tol;
    int max_iter, lmax, moment_mode, hoist_bc, grav_bc;
    int last_iter, last_status;
    double last_res;
};

static inline long cidx(const OrSolver *s, int i, int j, int k) {
    return ((long)i * s->ny + j) * s->nz + k;
}

/* ------------------------------------------------------------------ EOS -- */

/* hydro_solver_3d.c:162-175 */
static Prim c2p(Cons U, double gamma) {
    Prim P;
    P.d = omax(U.d, OR_RHO_FLOOR);
    P.u = U.mx / P.d;
    P.v = U.my / P.d;
    P.w = U.mz / P.d;
    double ke = 0.5 * P.d * (P.u * P.u + P.v * P.v + P.w * P.w);
    P.p = (gamma - 1.0) * (U.e - ke);
    P.p = omax(P.p, OR_P_FLOOR);
    return P;
}

/* hydro_solver_3d.c:177-189 */
static Cons p2c(Prim P, double gamma) {
    Cons U;
    U.d = P.d;
    U.mx = P.d * P.u;
    U.my = P.d * P.v;
    U.mz = P.d * P.w;
    double ke = 0.5 * P.d * (P.u * P.u + P.v * P.v + P.w * P.w);
    U.e = P.p / (gamma - 1.0) + ke;
    return U;
}

/* hydro_solver_3d.c:191-193 */
static double csound(double rho, double p, double gamma) {
    return sqrt(gamma * p / rho);
}

/* ----------------------------------------------------------------- HLLC -- */

/* Physical flux along direction dir (0=x,1=y,2=z); hydro_solver_3d.c:199-233.
 * vn is the velocity normal to the face. */
static Cons pflux(int dir, Cons U, Prim P) {
    Cons F;
    double vn = dir == 0 ? P.u : (dir == 1 ? P.v : P.w);
    double mn = dir == 0 ? U.mx : (dir == 1 ? U.my : U.mz);
    F.d = mn;
    F.mx = U.mx * vn;
    F.my = U.my * vn;
    F.mz = U.mz * vn;
    if (dir == 0) F.mx = U.mx * vn + P.p;
    if (dir == 1) F.my = U.my * vn + P.p;
    if (dir == 2) F.mz = U.mz * vn + P.p;
    F.e = (U.e + P.p) * vn;
    return F;
}

/* hllc_riemann_solver_{x,y,z}_3d, hydro_solver_3d.c:239-416.  The three
 * reference functions differ only in which velocity is normal; the star
 * state carries S* in the normal momentum and the side's own tangential
 * velocities. */
static Cons hllc(int dir, Cons UL, Cons UR, double gamma) {
    Prim PL = c2p(UL, gamma);
    Prim PR = c2p(UR, gamma);
    double cL = csound(PL.d, PL.p, gamma);
    double cR = csound(PR.d, PR.p, gamma);
    double uL = dir == 0 ? PL.u : (dir == 1 ? PL.v : PL.w);
    double uR = dir == 0 ? PR.u : (dir == 1 ? PR.v : PR.w);

    double SL = omin(uL - cL, uR - cR);
    double SR = omax(uL + cL, uR + cR);
    if (SL >= 0.0) return pflux(dir, UL, PL);
    if (SR <= 0.0) return pflux(dir, UR, PR);

    double num = PR.p - PL.p + PL.d * uL * (SL - uL) - PR.d * uR * (SR - uR);
    double den = PL.d * (SL - uL) - PR.d * (SR - uR);
    double Ss = num / den;

    /* choose the side by the sign of S* (hydro_solver_3d.c:265) */
    int left = (Ss >= 0.0);
    Cons Uk = left ? UL : UR;
    Prim Pk = left ? PL : PR;
    double Sk = left ? SL : SR;
    double uk = left ? uL : uR;

    double factor = (Sk - uk) / (Sk - Ss);
    Cons Us;
    Us.d = Pk.d * factor;
    Us.mx = Us.d * (dir == 0 ? Ss : Pk.u);
    Us.my = Us.d * (dir == 1 ? Ss : Pk.v);
    Us.mz = Us.d * (dir == 2 ? Ss : Pk.w);
    Us.e = Us.d * (Uk.e / Pk.d + (Ss - uk) * (Ss + Pk.p / (Pk.d * (Sk - uk))));

    Cons Fk = pflux(dir, Uk, Pk);
    Cons F;
    F.d = Fk.d + Sk * (Us.d - Uk.d);
    F.mx = Fk.mx + Sk * (Us.mx - Uk.mx);
    F.my = Fk.my + Sk * (Us.my - Uk.my);
    F.mz = Fk.mz + Sk * (Us.mz - Uk.mz);
    F.e = Fk.e + Sk * (Us.e - Uk.e);
    return F;
}

void or_cons_to_prim_batch(int n, const double *U, double gamma, double *P) {
    for (int c = 0; c < n; c++) {
        Cons u = {U[5 * c], U[5 * c + 1], U[5 * c + 2], U[5 * c + 3], U[5 * c + 4]};
        Prim p = c2p(u, gamma);
        P[5 * c] = p.d; P[5 * c + 1] = p.u; P[5 * c + 2] = p.v;
        P[5 * c + 3] = p.w; P[5 * c + 4] = p.p;
    }
}

void or_prim_to_cons_batch(int n, const double *P, double gamma, double *U) {
    for (int c = 0; c < n; c++) {
        Prim p = {P[5 * c], P[5 * c + 1], P[5 * c + 2], P[5 * c + 3], P[5 * c + 4]};
        Cons u = p2c(p, gamma);
        U[5 * c] = u.d; U[5 * c + 1] = u.mx; U[5 * c + 2] = u.my;
        U[5 * c + 3] = u.mz; U[5 * c + 4] = u.e;
    }
}

void or_hllc_batch(int dir, int n, const double *UL, const double *UR,
                   double gamma, double *F) {
    for (int c = 0; c < n; c++) {
        Cons l, r;
        memcpy(&l, UL + 5 * c, sizeof l);
        memcpy(&r, UR + 5 * c, sizeof r);
        Cons f = hllc(dir, l, r, gamma);
        memcpy(F + 5 * c, &f, sizeof f);
    }
}

/* ------------------------------------------------------------------- dt -- */

/* compute_dt_3d, hydro_solver_3d.c:422-450 */
static double dt_of(const Cons *U, long n, double dx, double dy, double dz,
                    double cfl, double gamma) {
    double ms = 1.0e-10;
    for (long c = 0; c < n; c++) {
        Prim P = c2p(U[c], gamma);
        double cs = csound(P.d, P.p, gamma);
        ms = omax(ms, fabs(P.u) + cs);
        ms = omax(ms, fabs(P.v) + cs);
        ms = omax(ms, fabs(P.w) + cs);
    }
    double tx = dx / ms, ty = dy / ms, tz = dz / ms;
    return cfl * omin(tx, omin(ty, tz));
}

double or_compute_dt(const double *U, int nx, int ny, int nz, double dx,
                     double dy, double dz, double cfl, double gamma) {
    return dt_of((const Cons *)U, (long)nx * ny * nz, dx, dy, dz, cfl, gamma);
}

/* ---------------------------------------------------- double-double sums -- */

typedef struct { double hi, lo; } DD;

static inline DD two_sum(double a, double b) {
    DD r;
    r.hi = a + b;
    double bb = r.hi - a;
    r.lo = (a - (r.hi - bb)) + (b - bb);
    return r;
}
static inline DD fast_two_sum(double a, double b) {
    DD r;
    r.hi = a + b;
    r.lo = b - (r.hi - a);
    return r;
}
/* Accurate double-double addition (Joldes, Muller, Popescu 2017, Alg. 6). */
static inline DD dd_add(DD x, DD y) {
    DD s = two_sum(x.hi, y.hi);
    DD t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    DD v = fast_two_sum(s.hi, c);
    double w = t.lo + v.lo;
    return fast_two_sum(v.hi, w);
}

/* Streaming pairwise tree: leaves pushed in index order; equal-size
 * subtrees merge as soon as they complete (binary counter); the leftover
 * stack is folded from the smallest subtree upward, which is the value of
 * the tree zero-padded to the next power of two. */
typedef struct {
    DD node[64];
    int level[64];
    int top;
} DDTree;

static void tree_init(DDTree *t) { t->top = 0; }
static void tree_push(DDTree *t, double v) {
    DD cur = {v, 0.0};
    int lev = 0;
    while (t->top > 0 && t->level[t->top - 1] == lev) {
        /* two leaves meet in an exact two-sum, later nodes in dd_add */
        cur = lev == 0 ? two_sum(t->node[t->top - 1].hi, v) : dd_add(t->node[t->top - 1], cur);
        t->top--;
        lev++;
    }
    t->node[t->top] = cur;
    t->level[t->top] = lev;
    t->top++;
}
static double tree_result(DDTree *t) {
    if (t->top == 0) return 0.0;
    DD cur = t->node[t->top - 1];
    for (int q = t->top - 2; q >= 0; q--) cur = dd_add(t->node[q], cur);
    return cur.hi;
}

double or_ddtree_sum(const double *v, long n) {
    DDTree t;
    tree_init(&t);
    for (long i = 0; i < n; i++) tree_push(&t, v[i]);
    return tree_result(&t);
}

/* --------------------------------------------------- solid harmonics -- */

/* Regular solid harmonics R_l^m = r^l P_l^m(cos th) e^{i m ph} (no
 * Condon-Shortley phase) as polynomials in x, y, z:
 *   R_m^m = (2m-1)!! (x + i y)^m,  R_{m+1}^m = (2m+1) z R_m^m,
 *   (l-m) R_l^m = (2l-1) z R_{l-1}^m - (l+m-1) r^2 R_{l-2}^m.
 * C = Re R, S = Im R.  (No reference counterpart: the reference stops at the
 * Cartesian quadrupole, gravity_solver_3d.c:196-262.) */
void or_solid_harmonics(int lmax, double x, double y, double z, double *C,
                        double *S) {
    int L = lmax + 1;
    double r2 = x * x + y * y + z * z;
    double cr = 1.0, ci = 0.0;
    for (int m = 0; m <= lmax; m++) {
        if (m > 0) {
            double f = (double)(2 * m - 1);
            double nr = f * (x * cr - y * ci);
            double ni = f * (x * ci + y * cr);
            cr = nr;
            ci = ni;
        }
        C[m * L + m] = cr;
        S[m * L + m] = ci;
        if (m + 1 <= lmax) {
            double f = (double)(2 * m + 1);
            C[(m + 1) * L + m] = f * z * cr;
            S[(m + 1) * L + m] = f * z * ci;
        }
        for (int l = m + 2; l <= lmax; l++) {
            double a = (double)(2 * l - 1), b = (double)(l + m - 1);
            double d = (double)(l - m);
            C[l * L + m] = (a * z * C[(l - 1) * L + m] - b * r2 * C[(l - 2) * L + m]) / d;
            S[l * L + m] = (a * z * S[(l - 1) * L + m] - b * r2 * S[(l - 2) * L + m]) / d;
        }
    }
}

static double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= (double)i;
    return f;
}
/* weight of the (l,m) term in the addition theorem for 1/|r-r'| */
static double mp_weight(int l, int m) {
    double w = fact(l - m) / fact(l + m);
    return m == 0 ? w : 2.0 * w;
}

/* ------------------------------------------------------------ moments -- */

/* Number of hi-order moment sums for a given lmax (A and B per l>=3, m<=l). */
static int n_hi(int lmax) {
    int n = 0;
    for (int l = 3; l <= lmax; l++) n += 2 * (l + 1);
    return n;
}

/* compute_mass_moments + compute_quadrupole_moment,
 * gravity_solver_3d.c:164-262, plus the l>=3 extension.  Two passes as in
 * the reference: the centre of mass first, then moments about it. */
static void moments(const OrSolver *s, const double *rho, double *out,
                    double *hi) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    const double dV = dx * dy * dz;
    const int tree = (s->moment_mode == OR_MOM_DDTREE);
    const int lmax = s->lmax;
    const int nh = n_hi(lmax);
    const int L = lmax + 1;

    /* pass 1: M, sum dm*x, sum dm*y, sum dm*z */
    double acc[4] = {0, 0, 0, 0};
    DDTree t4[4];
    for (int q = 0; q < 4; q++) tree_init(&t4[q]);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double r = rho[cidx(s, i, j, k)];
                double term[4] = {0, 0, 0, 0};
                int in = (r > 0.0);
                if (in) {
                    double x = s->xmin + (i + 0.5) * dx;
                    double y = s->ymin + (j + 0.5) * dy;
                    double z = s->zmin + (k + 0.5) * dz;
                    double dm = r * dV;
                    term[0] = dm; term[1] = dm * x; term[2] = dm * y; term[3] = dm * z;
                }
                if (tree) {
                    for (int q = 0; q < 4; q++) tree_push(&t4[q], term[q]);
                } else if (in) {
                    for (int q = 0; q < 4; q++) acc[q] += term[q];
                }
            }
    if (tree)
        for (int q = 0; q < 4; q++) acc[q] = tree_result(&t4[q]);
    double M = acc[0], xc, yc, zc;
    if (M > 0.0) {
        xc = acc[1] / M; yc = acc[2] / M; zc = acc[3] / M;
    } else {
        xc = 0.5 * (s->xmin + s->xmax);
        yc = 0.5 * (s->ymin + s->ymax);
        zc = 0.5 * (s->zmin + s->zmax);
    }
    out[0] = M; out[1] = xc; out[2] = yc; out[3] = zc;

    /* pass 2: Q_ij and the l>=3 moments about the centre of mass */
    int nq = 9 + nh;
    double *qa = calloc(nq, sizeof(double));
    DDTree *tq = malloc(sizeof(DDTree) * nq);
    double *term = malloc(sizeof(double) * nq);
    double C[(OR_LMAX_MAX + 1) * (OR_LMAX_MAX + 1)], S[(OR_LMAX_MAX + 1) * (OR_LMAX_MAX + 1)];
    for (int q = 0; q < nq; q++) tree_init(&tq[q]);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double r = rho[cidx(s, i, j, k)];
                int in = (r > 0.0);
                for (int q = 0; q < nq; q++) term[q] = 0.0;
                if (in) {
                    double x = s->xmin + (i + 0.5) * dx - xc;
                    double y = s->ymin + (j + 0.5) * dy - yc;
                    double z = s->zmin + (k + 0.5) * dz - zc;
                    double dm = r * dV;
                    double r2 = x * x + y * y + z * z;
                    term[0] = dm * (3.0 * x * x - r2);
                    term[1] = dm * (3.0 * x * y);
                    term[2] = dm * (3.0 * x * z);
                    term[3] = dm * (3.0 * y * x);
                    term[4] = dm * (3.0 * y * y - r2);
                    term[5] = dm * (3.0 * y * z);
                    term[6] = dm * (3.0 * z * x);
                    term[7] = dm * (3.0 * z * y);
                    term[8] = dm * (3.0 * z * z - r2);
                    if (nh) {
                        or_solid_harmonics(lmax, x, y, z, C, S);
                        int q = 9;
                        for (int l = 3; l <= lmax; l++)
                            for (int m = 0; m <= l; m++) {
                                term[q++] = dm * C[l * L + m];
                                term[q++] = dm * S[l * L + m];
                            }
                    }
                }
                if (tree) {
                    for (int q = 0; q < nq; q++) tree_push(&tq[q], term[q]);
                } else if (in) {
                    for (int q = 0; q < nq; q++) qa[q] += term[q];
                }
            }
    if (tree)
        for (int q = 0; q < nq; q++) qa[q] = tree_result(&tq[q]);
    for (int q = 0; q < 9; q++) out[4 + q] = qa[q];
    if (hi)
        for (int q = 0; q < nh; q++) hi[q] = qa[9 + q];
    free(qa);
    free(tq);
    free(term);
}

void or_moments(const OrSolver *s, const double *rho, double *out,
                double *out_hi) {
    moments(s, rho, out, out_hi);
}

/* Potential of the multipole expansion at a boundary cell centre;
 * gravity_solver_3d.c:341-360 for the monopole + quadrupole part. */
static double mp_phi(const OrSolver *s, const double *mom, const double *hi,
                     double x, double y, double z) {
    const double G = s->G;
    double M = mom[0];
    const double *Q = mom + 4;
    double rx = x - mom[1], ry = y - mom[2], rz = z - mom[3];
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    double ph = 0.0;
    if (r > 0.0) {
        ph = -G * M / r;
        double r5 = r * r * r * r * r;
        double Qr = Q[0] * rx * rx + Q[1] * rx * ry + Q[2] * rx * rz +
                    Q[3] * ry * rx + Q[4] * ry * ry + Q[5] * ry * rz +
                    Q[6] * rz * rx + Q[7] * rz * ry + Q[8] * rz * rz;
        double pq = -G * Qr / (2.0 * r5);
        ph += pq;
        if (s->lmax > 2) {
            int lmax = s->lmax, L = lmax + 1;
            double C[(OR_LMAX_MAX + 1) * (OR_LMAX_MAX + 1)], S[(OR_LMAX_MAX + 1) * (OR_LMAX_MAX + 1)];
            or_solid_harmonics(lmax, rx, ry, rz, C, S);
            double extra = 0.0, rp = r5;
            int q = 0;
            for (int l = 3; l <= lmax; l++) {
                rp = rp * r * r;
                double sl = 0.0;
                for (int m = 0; m <= l; m++) {
                    double A = hi[q++], B = hi[q++];
                    sl = sl + mp_weight(l, m) * (A * C[l * L + m] + B * S[l * L + m]);
                }
                extra = extra + sl / rp;
            }
            ph = ph + (-G) * extra;
        }
    }
    return ph;
}

/* apply_boundary_conditions, isolated branch, gravity_solver_3d.c:302-497:
 * x faces, then y faces, then z faces (nz > 1); each face uses its own
 * coordinate formula and later faces overwrite edges. */
static void boundary_isolated(const OrSolver *s, const double *mom,
                              const double *hi, double *phi) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            double y = s->ymin + (j + 0.5) * dy, z = s->zmin + (k + 0.5) * dz;
            phi[cidx(s, 0, j, k)] = mp_phi(s, mom, hi, s->xmin + 0.5 * dx, y, z);
            phi[cidx(s, nx - 1, j, k)] = mp_phi(s, mom, hi, s->xmax - 0.5 * dx, y, z);
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            double x = s->xmin + (i + 0.5) * dx, z = s->zmin + (k + 0.5) * dz;
            phi[cidx(s, i, 0, k)] = mp_phi(s, mom, hi, x, s->ymin + 0.5 * dy, z);
            phi[cidx(s, i, ny - 1, k)] = mp_phi(s, mom, hi, x, s->ymax - 0.5 * dy, z);
        }
    if (nz > 1)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                double x = s->xmin + (i + 0.5) * dx, y = s->ymin + (j + 0.5) * dy;
                phi[cidx(s, i, j, 0)] = mp_phi(s, mom, hi, x, y, s->zmin + 0.5 * dz);
                phi[cidx(s, i, j, nz - 1)] = mp_phi(s, mom, hi, x, y, s->zmax - 0.5 * dz);
            }
}

/* periodic branch, gravity_solver_3d.c:273-299 */
static void boundary_periodic(const OrSolver *s, double *phi) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) phi[cidx(s, 0, j, k)] = phi[cidx(s, nx - 1, j, k)];
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) phi[cidx(s, i, 0, k)] = phi[cidx(s, i, ny - 1, k)];
    if (nz > 1)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) phi[cidx(s, i, j, 0)] = phi[cidx(s, i, j, nz - 1)];
}

void or_boundary_phi(const OrSolver *s, const double *rho, double *phi) {
    double mom[13];
    double hi[128];
    moments(s, rho, mom, hi);
    boundary_isolated(s, mom, hi, phi);
}

/* ---------------------------------------------------------------- SOR -- */

/* gravity_solve_potential, gravity_solver_3d.c:510-747.  hoist_bc computes
 * the moments once per solve instead of every iteration: the density is
 * fixed during a solve, so the boundary values are identical either way. */
static int sor_solve(OrSolver *s, const double *rho) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx2_inv = 1.0 / (s->dx * s->dx);
    const double dy2_inv = 1.0 / (s->dy * s->dy);
    const double dz2_inv = 1.0 / (s->dz * s->dz);
    const double fpg = 4.0 * OR_PI * s->G;
    const double omega = s->omega;
    const int two_d = (nz == 1);
    const double factor = two_d ? 2.0 * (dx2_inv + dy2_inv)
                                : 2.0 * (dx2_inv + dy2_inv + dz2_inv);
    double *phi = s->phi;
    double mom[13], hi[128];
    int have_mom = 0;
    int iter;
    double maxres = 0.0;

    for (iter = 0; iter < s->max_iter; iter++) {
        maxres = 0.0;
        for (int color = 0; color < 2; color++) {
            if (two_d) {
                for (int i = 1; i < nx - 1; i++)
                    for (int j = 1; j < ny - 1; j++) {
                        if ((i + j) % 2 != color) continue;
                        long c = cidx(s, i, j, 0);
                        double old = phi[c];
                        double lap = (phi[cidx(s, i - 1, j, 0)] + phi[cidx(s, i + 1, j, 0)]) * dx2_inv +
                                     (phi[cidx(s, i, j - 1, 0)] + phi[cidx(s, i, j + 1, 0)]) * dy2_inv;
                        double rhs = fpg * rho[c];
                        double gs = (lap - rhs) / factor;
                        phi[c] = old + omega * (gs - old);
                        maxres = omax(maxres, fabs(gs - old));
                    }
            } else {
                for (int i = 1; i < nx - 1; i++)
                    for (int j = 1; j < ny - 1; j++)
                        for (int k = 1; k < nz - 1; k++) {
                            if ((i + j + k) % 2 != color) continue;
                            long c = cidx(s, i, j, k);
                            double old = phi[c];
                            double lap = (phi[c - (long)ny * nz] + phi[c + (long)ny * nz]) * dx2_inv +
                                         (phi[c - nz] + phi[c + nz]) * dy2_inv +
                                         (phi[c - 1] + phi[c + 1]) * dz2_inv;
                            double rhs = fpg * rho[c];
                            double gs = (lap - rhs) / factor;
                            phi[c] = old + omega * (gs - old);
                            maxres = omax(maxres, fabs(gs - old));
                        }
            }
        }
        /* gravity_solver_3d.c:712 */
        if (s->grav_bc == OR_GRAV_PERIODIC) {
            boundary_periodic(s, phi);
        } else if (s->grav_bc == OR_GRAV_ISOLATED) {
            if (!have_mom || !s->hoist_bc) {
                moments(s, rho, mom, hi);
                have_mom = 1;
            }
            boundary_isolated(s, mom, hi, phi);
        }
        if (maxres < s->tol) { /* gravity_solver_3d.c:715-727 */
            s->last_iter = iter + 1;
            s->last_res = maxres;
            return 0;
        }
    }
    s->last_iter = iter; /* gravity_solver_3d.c:730-746 */
    s->last_res = maxres;
    return -1;
}

/* gravity_compute_acceleration, gravity_solver_3d.c:753-899 */
static void accel(OrSolver *s) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    double *phi = s->phi, *gx = s->gx, *gy = s->gy, *gz = s->gz;
    if (nz == 1) {
        for (int i = 1; i < nx - 1; i++)
            for (int j = 1; j < ny - 1; j++) {
                long c = cidx(s, i, j, 0);
                gx[c] = -(phi[cidx(s, i + 1, j, 0)] - phi[cidx(s, i - 1, j, 0)]) / (2.0 * dx);
                gy[c] = -(phi[cidx(s, i, j + 1, 0)] - phi[cidx(s, i, j - 1, 0)]) / (2.0 * dy);
                gz[c] = 0.0;
            }
        for (int j = 0; j < ny; j++) {
            long a = cidx(s, 0, j, 0), b = cidx(s, nx - 1, j, 0);
            gx[a] = gy[a] = gz[a] = 0.0;
            gx[b] = gy[b] = gz[b] = 0.0;
        }
        for (int i = 0; i < nx; i++) {
            long a = cidx(s, i, 0, 0), b = cidx(s, i, ny - 1, 0);
            gx[a] = gy[a] = gz[a] = 0.0;
            gx[b] = gy[b] = gz[b] = 0.0;
        }
        return;
    }
    const long sx = (long)ny * nz, sy = nz;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                long c = cidx(s, i, j, k);
                gx[c] = -(phi[c + sx] - phi[c - sx]) / (2.0 * dx);
                gy[c] = -(phi[c + sy] - phi[c - sy]) / (2.0 * dy);
                gz[c] = -(phi[c + 1] - phi[c - 1]) / (2.0 * dz);
            }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            long a = cidx(s, 0, j, k), b = cidx(s, nx - 1, j, k);
            gx[a] = gx[a + sx]; gy[a] = gy[a + sx]; gz[a] = gz[a + sx];
            gx[b] = gx[b - sx]; gy[b] = gy[b - sx]; gz[b] = gz[b - sx];
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            long a = cidx(s, i, 0, k), b = cidx(s, i, ny - 1, k);
            gx[a] = gx[a + sy]; gy[a] = gy[a + sy]; gz[a] = gz[a + sy];
            gx[b] = gx[b - sy]; gy[b] = gy[b - sy]; gz[b] = gz[b - sy];
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            long a = cidx(s, i, j, 0), b = cidx(s, i, j, nz - 1);
            gx[a] = gx[a + 1]; gy[a] = gy[a + 1]; gz[a] = gz[a + 1];
            gx[b] = gx[b - 1]; gy[b] = gy[b - 1]; gz[b] = gz[b - 1];
        }
}

/* gravity_solve, gravity_solver_3d.c:905-911 */
int or_gravity_solve(OrSolver *s, const double *rho) {
    int st = sor_solve(s, rho);
    if (st == 0) accel(s);
    s->last_status = st;
    return st;
}

double or_rel_residual(const OrSolver *s, const double *rho,
                       const double *phi) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx2_inv = 1.0 / (s->dx * s->dx);
    const double dy2_inv = 1.0 / (s->dy * s->dy);
    const double dz2_inv = 1.0 / (s->dz * s->dz);
    const double fpg = 4.0 * OR_PI * s->G;
    const double factor = 2.0 * (dx2_inv + dy2_inv + dz2_inv);
    double rr = 0.0, ff = 0.0;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                long c = cidx(s, i, j, k);
                double lap = (phi[c - (long)ny * nz] + phi[c + (long)ny * nz]) * dx2_inv +
                             (phi[c - nz] + phi[c + nz]) * dy2_inv +
                             (phi[c - 1] + phi[c + 1]) * dz2_inv - factor * phi[c];
                double f = fpg * rho[c];
                double r = f - lap;
                rr += r * r;
                ff += f * f;
            }
    return ff > 0.0 ? sqrt(rr / ff) : sqrt(rr);
}

/* ------------------------------------------------------------- solver -- */

OrSolver *or_create(int nx, int ny, int nz, double xmin, double xmax,
                    double ymin, double ymax, double zmin, double zmax,
                    double G) {
    OrSolver *s = calloc(1, sizeof(OrSolver));
    long n = (long)nx * ny * nz;
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->xmin = xmin; s->xmax = xmax; s->ymin = ymin; s->ymax = ymax;
    s->zmin = zmin; s->zmax = zmax;
    /* hydro_solver_3d.c:55-62 */
    s->dx = (xmax - xmin) / nx; s->dy = (ymax - ymin) / ny; s->dz = (zmax - zmin) / nz;
    s->cfl = 0.4; s->dt = 0.0; s->gamma = 5.0 / 3.0; s->G = G;
    s->gravity = (G > 0.0);
    for (int q = 0; q < 6; q++) s->bc[q] = OR_BC_PERIODIC;
    s->U = calloc(n, sizeof(Cons));
    s->Fx = calloc(n, sizeof(Cons));
    s->Fy = calloc(n, sizeof(Cons));
    s->Fz = calloc(n, sizeof(Cons));
    s->hgx = calloc(n, sizeof(double));
    s->hgy = calloc(n, sizeof(double));
    s->hgz = calloc(n, sizeof(double));
    s->phi = calloc(n, sizeof(double));
    s->gx = calloc(n, sizeof(double));
    s->gy = calloc(n, sizeof(double));
    s->gz = calloc(n, sizeof(double));
    /* gravity_solver_3d.c:77-82; hydro switches to ISOLATED, :78 */
    s->omega = 1.5; s->max_iter = 10000; s->tol = 1.0e-6;
    s->lmax = 2; s->moment_mode = OR_MOM_SEQUENTIAL; s->hoist_bc = 1;
    s->grav_bc = OR_GRAV_ISOLATED;
    return s;
}

void or_destroy(OrSolver *s) {
    if (!s) return;
    free(s->U); free(s->Fx); free(s->Fy); free(s->Fz);
    free(s->hgx); free(s->hgy); free(s->hgz);
    free(s->phi); free(s->gx); free(s->gy); free(s->gz);
    free(s);
}

void or_set_hydro(OrSolver *s, double gamma, double cfl, const int bc[6]) {
    s->gamma = gamma;
    s->cfl = cfl;
    for (int q = 0; q < 6; q++) s->bc[q] = bc[q];
}

void or_set_gravity(OrSolver *s, double omega, int max_iter, double tol,
                    int lmax, int moment_mode, int hoist_bc, int grav_bc) {
    s->omega = omega; s->max_iter = max_iter; s->tol = tol;
    s->lmax = lmax < 2 ? 2 : (lmax > OR_LMAX_MAX ? OR_LMAX_MAX : lmax);
    s->moment_mode = moment_mode; s->hoist_bc = hoist_bc; s->grav_bc = grav_bc;
}

static long ncell(const OrSolver *s) { return (long)s->nx * s->ny * s->nz; }
void or_set_state(OrSolver *s, const double *U) { memcpy(s->U, U, ncell(s) * sizeof(Cons)); }
void or_get_state(const OrSolver *s, double *U) { memcpy(U, s->U, ncell(s) * sizeof(Cons)); }
void or_set_phi(OrSolver *s, const double *phi) { memcpy(s->phi, phi, ncell(s) * sizeof(double)); }
void or_get_phi(const OrSolver *s, double *phi) { memcpy(phi, s->phi, ncell(s) * sizeof(double)); }
void or_get_g(const OrSolver *s, double *gx, double *gy, double *gz) {
    memcpy(gx, s->hgx, ncell(s) * sizeof(double));
    memcpy(gy, s->hgy, ncell(s) * sizeof(double));
    memcpy(gz, s->hgz, ncell(s) * sizeof(double));
}
/* the gravity solver's own g (gravity_solver_3d.h: gx, gy, gz), which the
 * hydro side copies only after a converged solve */
void or_get_grav_g(const OrSolver *s, double *gx, double *gy, double *gz) {
    memcpy(gx, s->gx, ncell(s) * sizeof(double));
    memcpy(gy, s->gy, ncell(s) * sizeof(double));
    memcpy(gz, s->gz, ncell(s) * sizeof(double));
}
double or_get_dt(const OrSolver *s) { return s->dt; }
void or_get_diag(const OrSolver *s, int *iters, double *res, int *status) {
    if (iters) *iters = s->last_iter;
    if (res) *res = s->last_res;
    if (status) *status = s->last_status;
}

/* apply_boundary_conditions_3d, hydro_solver_3d.c:456-552 */
static void hydro_bc(OrSolver *s) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double g = s->gamma;
    Cons *U = s->U;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            long a = cidx(s, 0, j, k), b = cidx(s, nx - 1, j, k);
            if (s->bc[0] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[a], g); P.u = -P.u; U[a] = p2c(P, g);
            } else if (s->bc[0] == OR_BC_OUTFLOW) {
                U[a] = U[cidx(s, 1, j, k)];
            }
            if (s->bc[1] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[b], g); P.u = -P.u; U[b] = p2c(P, g);
            } else if (s->bc[1] == OR_BC_OUTFLOW) {
                U[b] = U[cidx(s, nx - 2, j, k)];
            }
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            long a = cidx(s, i, 0, k), b = cidx(s, i, ny - 1, k);
            if (s->bc[2] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[a], g); P.v = -P.v; U[a] = p2c(P, g);
            } else if (s->bc[2] == OR_BC_OUTFLOW) {
                U[a] = U[cidx(s, i, 1, k)];
            }
            if (s->bc[3] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[b], g); P.v = -P.v; U[b] = p2c(P, g);
            } else if (s->bc[3] == OR_BC_OUTFLOW) {
                U[b] = U[cidx(s, i, ny - 2, k)];
            }
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            long a = cidx(s, i, j, 0), b = cidx(s, i, j, nz - 1);
            if (s->bc[4] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[a], g); P.w = -P.w; U[a] = p2c(P, g);
            } else if (s->bc[4] == OR_BC_OUTFLOW) {
                U[a] = U[cidx(s, i, j, 1)];
            }
            if (s->bc[5] == OR_BC_REFLECTIVE) {
                Prim P = c2p(U[b], g); P.w = -P.w; U[b] = p2c(P, g);
            } else if (s->bc[5] == OR_BC_OUTFLOW) {
                U[b] = U[cidx(s, i, j, nz - 2)];
            }
        }
}

/* godunov_step_3d, hydro_solver_3d.c:637-811 (with solve_gravity_field
 * :558-591 and apply_gravity_source_terms :597-631). */
int or_step(OrSolver *s) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const long n = ncell(s);
    const double g = s->gamma;
    Cons *U = s->U;
    int gst = 0;

    if (s->gravity) {
        double *rho = malloc(n * sizeof(double));
        for (long c = 0; c < n; c++) rho[c] = U[c].d;
        gst = or_gravity_solve(s, rho);
        if (gst != 0) fprintf(stderr, "oracle: gravity solver did not converge\n");
        memcpy(s->hgx, s->gx, n * sizeof(double));
        memcpy(s->hgy, s->gy, n * sizeof(double));
        memcpy(s->hgz, s->gz, n * sizeof(double));
        free(rho);
    }

    s->dt = dt_of(U, n, s->dx, s->dy, s->dz, s->cfl, g);
    const double dt = s->dt, dx = s->dx, dy = s->dy, dz = s->dz;

    for (int i = 0; i < nx - 1; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                s->Fx[cidx(s, i, j, k)] = hllc(0, U[cidx(s, i, j, k)], U[cidx(s, i + 1, j, k)], g);
    if (s->bc[0] == OR_BC_PERIODIC && s->bc[1] == OR_BC_PERIODIC)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                s->Fx[cidx(s, nx - 1, j, k)] = hllc(0, U[cidx(s, nx - 1, j, k)], U[cidx(s, 0, j, k)], g);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny - 1; j++)
            for (int k = 0; k < nz; k++)
                s->Fy[cidx(s, i, j, k)] = hllc(1, U[cidx(s, i, j, k)], U[cidx(s, i, j + 1, k)], g);
    if (s->bc[2] == OR_BC_PERIODIC && s->bc[3] == OR_BC_PERIODIC)
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++)
                s->Fy[cidx(s, i, ny - 1, k)] = hllc(1, U[cidx(s, i, ny - 1, k)], U[cidx(s, i, 0, k)], g);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz - 1; k++)
                s->Fz[cidx(s, i, j, k)] = hllc(2, U[cidx(s, i, j, k)], U[cidx(s, i, j, k + 1)], g);
    if (s->bc[4] == OR_BC_PERIODIC && s->bc[5] == OR_BC_PERIODIC)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                s->Fz[cidx(s, i, j, nz - 1)] = hllc(2, U[cidx(s, i, j, nz - 1)], U[cidx(s, i, j, 0)], g);

    /* conservative update + floors, hydro_solver_3d.c:729-771 */
    const double ax = dt / dx, ay = dt / dy, az = dt / dz;
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                long c = cidx(s, i, j, k);
                const Cons *X0 = &s->Fx[c - (long)ny * nz], *X1 = &s->Fx[c];
                const Cons *Y0 = &s->Fy[c - nz], *Y1 = &s->Fy[c];
                const Cons *Z0 = &s->Fz[c - 1], *Z1 = &s->Fz[c];
                Cons d;
                d.d = -ax * (X1->d - X0->d) - ay * (Y1->d - Y0->d) - az * (Z1->d - Z0->d);
                d.mx = -ax * (X1->mx - X0->mx) - ay * (Y1->mx - Y0->mx) - az * (Z1->mx - Z0->mx);
                d.my = -ax * (X1->my - X0->my) - ay * (Y1->my - Y0->my) - az * (Z1->my - Z0->my);
                d.mz = -ax * (X1->mz - X0->mz) - ay * (Y1->mz - Y0->mz) - az * (Z1->mz - Z0->mz);
                d.e = -ax * (X1->e - X0->e) - ay * (Y1->e - Y0->e) - az * (Z1->e - Z0->e);
                U[c].d += d.d; U[c].mx += d.mx; U[c].my += d.my; U[c].mz += d.mz; U[c].e += d.e;
                U[c].d = omax(U[c].d, OR_RHO_FLOOR);
                Prim P = c2p(U[c], g);
                P.p = omax(P.p, OR_P_FLOOR);
                U[c] = p2c(P, g);
            }

    if (s->gravity) { /* hydro_solver_3d.c:597-631, all cells */
        for (long c = 0; c < n; c++) {
            double r = U[c].d;
            double u = U[c].mx / r, v = U[c].my / r, w = U[c].mz / r;
            double gx = s->hgx[c], gy = s->hgy[c], gz = s->hgz[c];
            U[c].mx += dt * r * gx;
            U[c].my += dt * r * gy;
            U[c].mz += dt * r * gz;
            double work = r * (u * gx + v * gy + w * gz);
            U[c].e += dt * work;
        }
    }

    /* periodic copies, hydro_solver_3d.c:779-807 */
    if (s->bc[0] == OR_BC_PERIODIC && s->bc[1] == OR_BC_PERIODIC)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) U[cidx(s, 0, j, k)] = U[cidx(s, nx - 1, j, k)];
    if (s->bc[2] == OR_BC_PERIODIC && s->bc[3] == OR_BC_PERIODIC)
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) U[cidx(s, i, 0, k)] = U[cidx(s, i, ny - 1, k)];
    if (s->bc[4] == OR_BC_PERIODIC && s->bc[5] == OR_BC_PERIODIC)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) U[cidx(s, i, j, 0)] = U[cidx(s, i, j, nz - 1)];
    hydro_bc(s);
    return gst;
}

/* run_simulation_3d loop body without output, hydro_solver_3d.c:2257-2260 */
int or_run(OrSolver *s, double t_end, int max_steps, double *t_out,
           double *dt_log, int *iter_log) {
    double t = 0.0;
    int steps = 0;
    while (t < t_end) {
        or_step(s);
        t += s->dt;
        if (dt_log) dt_log[steps] = s->dt;
        if (iter_log) iter_log[steps] = s->gravity ? s->last_iter : 0;
        steps++;
        if (max_steps > 0 && steps >= max_steps) break;
    }
    if (t_out) *t_out = t;
    return steps;
}
