// gravity_api.hip -- the gravity_solver_3d.h API on the GPU.
//
// Lifecycle, parameters and diagnostics follow gravity_solver_3d.c:49-155;
// the solve follows gravity_solver_3d.c:510-747 and :753-911 with the
// potential, density and acceleration resident on the device.  The device
// potential is authoritative: the host arrays phi/gx/gy/gz are refreshed
// when an API call that promises them returns, and gravity_sync_to_device
// pushes a caller-edited phi back.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/hydro_mi355x.h"
#include "../../include/input_deck.h"
#include "host_common.hpp"

using hyd::GridDims;
using hyd::SorCoef;
using hyd::SorCtl;

hipStream_t hyd_stream();

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI of glibc
constexpr int kW = hyd::MP_LMAX_MAX + 1;

double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= (double)i;
    return f;
}

template <class T>
int dev_alloc(T** p, long count) {
    HYD_CHECK(hipMalloc((void**)p, sizeof(T) * (size_t)(count > 0 ? count : 1)));
    HYD_CHECK(hipMemsetAsync(*p, 0, sizeof(T) * (size_t)(count > 0 ? count : 1), hyd_stream()));
    return 0;
}

// Whether the moments are summed in the reference's order: asked for, or
// by default for plain FP64 SOR (the defect-correction modes keep the
// compensated tree).
static bool sequential_moments(const GravitySolver3D* s) {
    return s->moment_order == MOMENT_ORDER_SEQUENTIAL ||
           (s->moment_order == MOMENT_ORDER_AUTO && s->poisson_mode == POISSON_FP64_SOR);
}

// The multipole moments of rho (device) into d->mom, in the solver's order.
int moments_dev(GravitySolver3D* s, const double* d_rho) {
    GravDevice* d = s->dev;
    if (!sequential_moments(s)) {
        HYD_CHECK(hyd::launch_moments(d->g, d_rho, s->multipole_lmax, d->mom_scratch, d->mom,
                                      d->stream));
        return 0;
    }
    const long len = hyd::moments_seq_len(d->n, s->multipole_lmax);
    if (len > d->mom_seq_len) {
        if (d->mom_seq) HYD_CHECK(hipFree(d->mom_seq));
        d->mom_seq = nullptr;
        d->mom_seq_len = 0;
        if (dev_alloc(&d->mom_seq, len)) return -2;
        d->mom_seq_len = len;
    }
    HYD_CHECK(hyd::launch_moments_sequential(d->g, d_rho, s->multipole_lmax, d->seq_method,
                                             d->mom_seq, d->mom, d->stream));
    return 0;
}

void dev_free(GravDevice* d) {
    if (!d) return;
    void* ptrs[] = {d->phi_r, d->phi_b, d->rhs_r, d->rhs_b, d->gx,  d->gy, d->gz,
                    d->rho_tmp, d->nat_tmp, d->mom, d->mom_scratch, d->mom_seq, d->wlm, d->ctl, d->acc, d->res_part,
                    d->e_r, d->e_b, d->r_r, d->r_b, d->guess_r, d->guess_b,
};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto* v : {&d->mg_u, &d->mg_f, &d->mg_q})
        for (float* p : *v)
            if (p) (void)hipFree(p);
    if (d->mg_qr) (void)hipFree(d->mg_qr);
    if (d->mg_qb) (void)hipFree(d->mg_qb);
    if (d->mg_tail_dev) (void)hipFree(d->mg_tail_dev);
    if (d->ctl_host) (void)hipHostFree(d->ctl_host);
    if (d->acc_host) (void)hipHostFree(d->acc_host);
    if (d->ev0) (void)hipEventDestroy(d->ev0);
    if (d->ev1) (void)hipEventDestroy(d->ev1);
    delete d;
}

int dev_create(GravitySolver3D* s) {
    GravDevice* d = new GravDevice();
    s->dev = d;
    GridDims& g = d->g;
    g.nx = s->nx; g.ny = s->ny; g.nz = s->nz; g.nzh = (s->nz + 1) / 2;
    g.xmin = s->xmin; g.xmax = s->xmax; g.ymin = s->ymin; g.ymax = s->ymax;
    g.zmin = s->zmin; g.zmax = s->zmax;
    g.dx = s->dx; g.dy = s->dy; g.dz = s->dz;
    d->n = (long)s->nx * s->ny * s->nz;
    d->nsplit = (long)s->nx * s->ny * g.nzh;
    d->two_d = (s->nz == 1);
    d->stream = hyd_stream();
    if (!d->stream) return -2;
    if (const char* e = getenv("HYDRO_MI355X_MG_TAIL")) d->mg_use_tail = atoi(e) != 0;
    if (const char* e = getenv("HYDRO_MI355X_SEQSUM"))
        d->seq_method = strcmp(e, "chain") == 0     ? hyd::SEQSUM_CHAIN
                        : strcmp(e, "checked") == 0 ? hyd::SEQSUM_PEAKS_CHECKED
                                                    : hyd::SEQSUM_PEAKS;
    if (const char* e = getenv("HYDRO_MI355X_MG_SMOOTH")) {
        int a = 0, b = 0;
        float w = 0.0f;
        if (sscanf(e, "%d,%d,%f", &a, &b, &w) == 3 && a >= 0 && b >= 0 && a + b >= 1 &&
            w > 0.0f && w < 2.0f) {
            d->mg_nu1 = a;
            d->mg_nu2 = b;
            d->mg_omega = w;
        }
    }
    if (dev_alloc(&d->phi_r, d->nsplit) || dev_alloc(&d->phi_b, d->nsplit) ||
        dev_alloc(&d->rhs_r, d->nsplit) || dev_alloc(&d->rhs_b, d->nsplit) ||
        dev_alloc(&d->gx, d->n) || dev_alloc(&d->gy, d->n) || dev_alloc(&d->gz, d->n) ||
        dev_alloc(&d->mom, 13 + hyd::mp_n_hi(hyd::MP_LMAX_MAX)) || dev_alloc(&d->wlm, kW * kW) ||
        dev_alloc(&d->ctl, 1) || dev_alloc(&d->acc, 2) ||
        dev_alloc(&d->res_part, hyd::residual_scratch_len(d->g)))
        return -2;
    d->mom_scratch_len = hyd::moments_scratch_len(d->n, hyd::MP_LMAX_MAX);
    if (dev_alloc(&d->mom_scratch, d->mom_scratch_len)) return -2;
    HYD_CHECK(hipHostMalloc((void**)&d->ctl_host, sizeof(SorCtl), hipHostMallocDefault));
    HYD_CHECK(hipHostMalloc((void**)&d->acc_host, 2 * sizeof(double), hipHostMallocDefault));
    HYD_CHECK(hipEventCreate(&d->ev0));
    HYD_CHECK(hipEventCreate(&d->ev1));
    // weight of the (l,m) term of the addition theorem for 1/|r - r'|
    double w[kW * kW];
    for (int l = 0; l < kW; ++l)
        for (int m = 0; m < kW; ++m) {
            double v = 0.0;
            if (m <= l) {
                v = fact(l - m) / fact(l + m);
                if (m > 0) v = 2.0 * v;
            }
            w[l * kW + m] = v;
        }
    HYD_CHECK(hipMemcpyAsync(d->wlm, w, sizeof w, hipMemcpyHostToDevice, d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

SorCoef coefficients(const GravitySolver3D* s) {
    // gravity_solver_3d.c:519-545
    SorCoef c;
    const double dx2 = s->dx * s->dx, dy2 = s->dy * s->dy, dz2 = s->dz * s->dz;
    c.dx2i = 1.0 / dx2;
    c.dy2i = 1.0 / dy2;
    c.dz2i = 1.0 / dz2;
    c.factor = s->nz == 1 ? 2.0 * (c.dx2i + c.dy2i)  // gravity_solver_3d.c:539-545
                          : 2.0 * (c.dx2i + c.dy2i + c.dz2i);
    c.omega = s->omega;
    return c;
}

double bitsd(unsigned long long b) {
    double v;
    memcpy(&v, &b, sizeof v);
    return v;
}

// Next batch of iterations from the observed convergence rate: enough to
// reach the tolerance if the rate holds, so that one host round trip usually
// ends the solve; extra launches after convergence exit at once.
int next_batch(double r1, double r2, double tol, int prev) {
    if (!(r1 > 0.0) || !(r2 > 0.0)) return std::min(2 * prev, 1024);
    const double rate = r1 / r2;
    if (!(rate < 1.0) || !(rate > 0.0)) return std::min(2 * prev, 1024);
    const double est = log(tol / r1) / log(rate);
    int b = (int)ceil(est * 1.02) + 2;
    return std::max(4, std::min(b, 4096));
}

int rel_residual_dev(GravitySolver3D* s, float* rr, float* rb, double* rel) {
    GravDevice* d = s->dev;
    SorCoef c = coefficients(s);
    HYD_CHECK(hyd::launch_residual(d->g, c, d->phi_r, d->phi_b, d->rhs_r, d->rhs_b, rr, rb,
                                   d->res_part, d->acc,
                                   d->stream));
    HYD_CHECK(hipMemcpyAsync(d->acc_host, d->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    const double rr2 = d->acc_host[0], ff = d->acc_host[1];
    *rel = ff > 0.0 ? sqrt(rr2 / ff) : sqrt(rr2);
    return 0;
}

int apply_boundary(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    const int bc = (s->bc_type == GRAV_BC_PERIODIC) ? 0 : (s->bc_type == GRAV_BC_ISOLATED ? 1 : 2);
    if (d->two_d) {
        HYD_CHECK(hyd::launch_sor2d_boundary(d->g, bc, s->G, s->multipole_lmax, d->mom, d->wlm,
                                             d->phi_r, d->stream));
        return 0;
    }
    HYD_CHECK(hyd::launch_sor_boundary(d->g, bc, s->G, s->multipole_lmax, d->mom, d->wlm, d->phi_r,
                                       d->phi_b, d->stream));
    return 0;
}

// Red-black SOR in FP64 with the reference's iteration structure
// (gravity_solver_3d.c:566-746): red, black, boundary, check.  The boundary
// values depend only on the density, so they are written once, after the
// first iteration, exactly where the reference first writes them.  Each
// iteration is two colour sweeps in place; under the max-update rule a
// one-wavefront check before each iteration evaluates the stop rule on the
// device, and sweeps past convergence return at their first instruction, so
// the host reads the state back only once per batch of iterations.
int sor_fp64(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    const SorCoef c = coefficients(s);
    const bool rel = (s->criterion == SOR_CRIT_REL_RESID);
    const double tol = s->tolerance;
    const double stol = rel ? -1.0 : tol;  // device-side stop rule only for max-update
    HYD_CHECK(hyd::launch_sor_reset(d->ctl, d->stream));
    if (s->max_iterations <= 0) {
        s->last_iterations = 0;
        s->last_residual = 0.0;
        return -1;
    }
    const double four_pi_G = 4.0 * kPi * s->G;
    auto iteration = [&](int it) -> int {
        // the stop rule of iteration it, evaluated once for both sweeps
        if (!rel) HYD_CHECK(hyd::launch_sor_check(d->ctl, it, tol, d->stream));
        if (d->two_d) {  // gravity_solver_3d.c:569-634, rho from the caller's array
            for (int colour = 0; colour < 2; ++colour)
                HYD_CHECK(hyd::launch_sor2d_sweep(d->g, c, colour, it, stol, four_pi_G, d->phi_r,
                                                  d->rho_2d, d->ctl, d->stream));
            return 0;
        }
        HYD_CHECK(hyd::launch_sor_sweep(d->g, c, 0, it, stol, d->phi_r, d->phi_b, d->rhs_r,
                                        d->rhs_b, d->ctl, d->stream));
        HYD_CHECK(hyd::launch_sor_sweep(d->g, c, 1, it, stol, d->phi_r, d->phi_b, d->rhs_r,
                                        d->rhs_b, d->ctl, d->stream));
        return 0;
    };
    // iteration 0, then the boundary values
    if (iteration(0)) return -2;
    if (apply_boundary(s)) return -2;
    int it = 1;
    // Under the max-update rule a warm start usually needs about as many
    // iterations as the last solve, and iterations past convergence return
    // at once: one batch of that size ends most solves with a single host
    // round trip.  (The relative-residual rule stops only between batches.)
    // After a long solve (a cold start) the next, warm one is far shorter:
    // it starts from a batch of 16 and grows, since every launch past
    // convergence still costs some microseconds.
    int batch = 8;
    if (!rel && s->last_iterations > 1) batch = s->last_iterations <= 128 ? s->last_iterations : 16;
    double rel_prev = -1.0;
    int it_prev = 0;
    while (true) {
        const int nb = std::min(batch, s->max_iterations - it);
        for (int q = 0; q < nb; ++q, ++it)
            if (iteration(it)) return -2;
        if (rel) {
            double r;
            if (rel_residual_dev(s, nullptr, nullptr, &r)) return -2;
            s->last_rel_residual = r;
            if (r <= tol || it >= s->max_iterations) {
                s->last_iterations = it;
                s->last_residual = r;
                return r <= tol ? 0 : -1;
            }
            int b = batch;
            if (rel_prev > 0.0 && r < rel_prev) {
                const double rate = pow(r / rel_prev, 1.0 / (it - it_prev));
                b = (int)ceil(log(tol / r) / log(rate) * 1.02) + 1;
            } else {
                b = 2 * batch;
            }
            rel_prev = r;
            it_prev = it;
            batch = std::max(8, std::min(b, 4096));
            continue;
        }
        // max-update criterion: the stop check of iteration `it`, then read back
        HYD_CHECK(hyd::launch_sor_check(d->ctl, it, tol, d->stream));
        HYD_CHECK(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(SorCtl), hipMemcpyDeviceToHost,
                                 d->stream));
        HYD_CHECK(hipStreamSynchronize(d->stream));
        const SorCtl& h = *d->ctl_host;
        if (h.converged) {
            s->last_iterations = h.iters;  // first iteration that did not run
            s->last_residual = h.last_res;
            return 0;
        }
        const double r1 = bitsd(hyd::sor_slot_max(h, (it - 1) & 3));
        if (it >= s->max_iterations) {
            s->last_iterations = it;
            s->last_residual = r1;
            return -1;
        }
        const double r2 = it >= 2 ? bitsd(hyd::sor_slot_max(h, (it - 2) & 3)) : 0.0;
        batch = next_batch(r1, r2, tol, batch);
    }
}

int ensure_dc_buffers(GravDevice* d) {
    if (!d->e_r && (dev_alloc(&d->e_r, d->nsplit) || dev_alloc(&d->e_b, d->nsplit) ||
                    dev_alloc(&d->r_r, d->nsplit) || dev_alloc(&d->r_b, d->nsplit)))
        return -2;
    return 0;
}

// Boundary values and residual r (FP32, split) of the starting potential,
// unless the start choice left them in place.
int start_residual(GravitySolver3D* s, double* rel) {
    GravDevice* d = s->dev;
    if (d->start_residual_ready) {
        d->start_residual_ready = false;
        *rel = d->start_rel;
        return 0;
    }
    if (apply_boundary(s)) return -2;
    return rel_residual_dev(s, d->r_r, d->r_b, rel);
}

// Defect correction: outer FP64 residual r = f - L phi, inner FP32
// red-black sweeps on L e = r with e = 0 on the boundary, phi += e in FP64.
// Stops on the relative residual ||r|| / ||f|| <= tolerance.
int sor_fp32_dc(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    if (ensure_dc_buffers(d)) return -2;
    const SorCoef c = coefficients(s);
    double rel;
    if (start_residual(s, &rel)) return -2;
    int sweeps = 0, k = 64;
    const double tol = s->tolerance;
    for (int outer = 0; outer < 64 && rel > tol && sweeps < s->max_iterations; ++outer) {
        const int nk = std::min(k, s->max_iterations - sweeps);
        for (int q = 0; q < nk; ++q) {
            HYD_CHECK(hyd::launch_sor_sweep_f32(d->g, c, 0, d->e_r, d->e_b, d->r_r, d->r_b,
                                                d->stream));
            HYD_CHECK(hyd::launch_sor_sweep_f32(d->g, c, 1, d->e_r, d->e_b, d->r_r, d->r_b,
                                                d->stream));
        }
        sweeps += nk;
        HYD_CHECK(hyd::launch_correct(d->g, d->phi_r, d->phi_b, d->e_r, d->e_b, d->stream));
        double rnew;
        if (rel_residual_dev(s, d->r_r, d->r_b, &rnew)) return -2;
        // per-sweep contraction of this round -> sweeps for the next round,
        // aiming at a 1e-4 reduction (well inside FP32 reach) or the target
        const double red = rnew / rel;
        if (red > 0.0 && red < 1.0) {
            const double rate = pow(red, 1.0 / nk);
            const double want = std::max(1e-4, 0.5 * tol / rnew);
            k = (int)ceil(log(want) / log(rate));
            k = std::max(8, std::min(k, 8192));
        } else {
            k = std::min(2 * k, 8192);
        }
        rel = rnew;
    }
    s->last_iterations = sweeps;
    s->last_residual = rel;
    s->last_rel_residual = rel;
    return rel <= tol ? 0 : -1;
}

bool mg_supported(const GravitySolver3D* s) {
    return s->nx % 2 == 0 && s->ny % 2 == 0 && s->nz % 2 == 0 && s->nx >= 8 && s->ny >= 8 &&
           s->nz >= 8;
}

// The coarse levels: level l has n / 2^l + 1 points per direction at fine
// cells 0 and c 2^l - 1 (c >= 1), halving while every direction has an even
// interval count and at least 5 points.  Per direction, with H = 2^l fine
// cells: the uniform coefficient 1 / (H h)^2; at index 1 (first interval
// (H - 1) h) the three-point weights 2 / ((hl + hr) hl) and 2 / ((hl + hr) hr);
// the interpolation weight (H - 1) / (2H - 1) of the next level's point 1 at
// index 1; and, with Hf = H / 2 the finer spacing, the restriction weights
// (Hf - 1, 2 Hf, Hf) / (4 Hf - 1) of the finer points around point 1 (the
// transpose of interpolation scaled by dual-cell lengths).
hyd::MgAxis mg_axis(double h, int l) {
    const double H = (double)(1 << l), Hf = H / 2.0;
    const double hl = (H - 1.0) * h, hr = H * h, sc = 2.0 / (hl + hr);
    hyd::MgAxis A;
    A.a = (float)(1.0 / (hr * hr));
    A.am1 = (float)(sc / hl);
    A.ap1 = (float)(sc / hr);
    A.pw1 = (float)((H - 1.0) / (2.0 * H - 1.0));
    A.r1[0] = (float)((Hf - 1.0) / (4.0 * Hf - 1.0));
    A.r1[1] = (float)(2.0 * Hf / (4.0 * Hf - 1.0));
    A.r1[2] = (float)(Hf / (4.0 * Hf - 1.0));
    return A;
}

int ensure_mg(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    if (!d->mg.empty()) return 0;
    hyd::MgLevel L;
    L.nx = s->nx / 2 + 1;
    L.ny = s->ny / 2 + 1;
    L.nz = s->nz / 2 + 1;
    for (int l = 1;; ++l) {
        L.ax[0] = mg_axis(s->dx, l);
        L.ax[1] = mg_axis(s->dy, l);
        L.ax[2] = mg_axis(s->dz, l);
        d->mg.push_back(L);
        const bool halve = (L.nx - 1) % 2 == 0 && (L.ny - 1) % 2 == 0 && (L.nz - 1) % 2 == 0 &&
                           std::min(L.nx, std::min(L.ny, L.nz)) >= 5;
        if (!halve) break;
        L.nx = (L.nx - 1) / 2 + 1;
        L.ny = (L.ny - 1) / 2 + 1;
        L.nz = (L.nz - 1) / 2 + 1;
    }
    for (const hyd::MgLevel& l : d->mg) {
        const long n = (long)l.nx * l.ny * l.nz;
        float *u = nullptr, *f = nullptr, *q = nullptr;
        if (dev_alloc(&u, n) || dev_alloc(&f, n) || dev_alloc(&q, n)) return -2;
        d->mg_u.push_back(u);
        d->mg_f.push_back(f);
        d->mg_q.push_back(q);
    }
    if (dev_alloc(&d->mg_qr, d->nsplit) || dev_alloc(&d->mg_qb, d->nsplit)) return -2;
    // the tail: the longest run of coarsest levels whose first level has at
    // most kTailPoints points and whose u, f, q fit the workgroup's LDS
    constexpr long kTailPoints = 6000;
    const int nl = (int)d->mg.size();
    d->mg_tail_first = -1;
    long floats = 0;
    for (int l = nl - 1; l >= 0 && d->mg_use_tail && nl - l <= hyd::MG_TAIL_MAX; --l) {
        const long pts = (long)d->mg[l].nx * d->mg[l].ny * d->mg[l].nz;
        if (pts > kTailPoints || sizeof(float) * (size_t)(floats + 3 * pts) > hyd::MG_TAIL_LDS_MAX)
            break;
        floats += 3 * pts;
        d->mg_tail_first = l;
    }
    if (d->mg_tail_first >= 0 && d->mg_tail_first < nl - 1) {
        hyd::MgTail tab{};
        tab.count = nl - d->mg_tail_first;
        tab.nu1 = d->mg_nu1;
        tab.nu2 = d->mg_nu2;
        tab.omega = d->mg_omega;
        int off = 0;
        for (int q = 0; q < tab.count; ++q) {
            const hyd::MgLevel& L = d->mg[d->mg_tail_first + q];
            const int pts = L.nx * L.ny * L.nz;
            tab.lv[q] = L;
            tab.off_u[q] = off;
            tab.off_f[q] = off + pts;
            tab.off_q[q] = off + 2 * pts;
            off += 3 * pts;
        }
        d->mg_tail_floats = off;
        if (dev_alloc(&d->mg_tail_dev, 1)) return -2;
        HYD_CHECK(hipMemcpyAsync(d->mg_tail_dev, &tab, sizeof tab, hipMemcpyHostToDevice,
                                 d->stream));
        HYD_CHECK(hipStreamSynchronize(d->stream));
    } else {
        d->mg_tail_first = -1;  // a single level gains nothing from the tail
    }
    return 0;
}

int mg_smooth(GravDevice* d, size_t l, int sweeps, float omega) {
    for (int q = 0; q < sweeps; ++q) {
        HYD_CHECK(hyd::launch_mg_smooth(d->mg[l], 0, omega, d->mg_u[l], d->mg_f[l], d->stream));
        HYD_CHECK(hyd::launch_mg_smooth(d->mg[l], 1, omega, d->mg_u[l], d->mg_f[l], d->stream));
    }
    return 0;
}

// V-cycle from coarse level l (index into d->mg) down, on u_l with f_l.
int mg_coarse_cycle(GravDevice* d, size_t l) {
    if ((int)l == d->mg_tail_first) {
        HYD_CHECK(hyd::launch_mg_tail(d->mg_tail_dev, d->mg_tail_floats, d->mg_f[l], d->mg_u[l],
                                      d->stream));
        return 0;
    }
    const hyd::MgLevel& L = d->mg[l];
    if (l + 1 == d->mg.size())  // coarsest: Gauss-Seidel sweeps
        return mg_smooth(d, l, std::max(L.nx, std::max(L.ny, L.nz)), 1.0f);
    if (mg_smooth(d, l, d->mg_nu1, d->mg_omega)) return -2;
    HYD_CHECK(hyd::launch_mg_residual(L, d->mg_u[l], d->mg_f[l], d->mg_q[l], d->stream));
    HYD_CHECK(hyd::launch_mg_restrict(L, d->mg_q[l], d->mg[l + 1], d->mg_f[l + 1],
                                      d->mg_u[l + 1], d->stream));
    if (mg_coarse_cycle(d, l + 1)) return -2;
    HYD_CHECK(hyd::launch_mg_prolong(d->mg[l + 1], d->mg_u[l + 1], L, d->mg_u[l], d->stream));
    return mg_smooth(d, l, d->mg_nu2, d->mg_omega);
}

// One V-cycle on the solver's grid: L e = r (FP32, split layout), e += ...
int mg_vcycle(GravitySolver3D* s, const SorCoef& c) {
    GravDevice* d = s->dev;
    SorCoef cs = c;
    cs.omega = d->mg_omega;
    auto sweep = [&](int n) -> int {
        for (int q = 0; q < n; ++q) {
            HYD_CHECK(hyd::launch_sor_sweep_f32(d->g, cs, 0, d->e_r, d->e_b, d->r_r, d->r_b,
                                                d->stream));
            HYD_CHECK(hyd::launch_sor_sweep_f32(d->g, cs, 1, d->e_r, d->e_b, d->r_r, d->r_b,
                                                d->stream));
        }
        return 0;
    };
    if (sweep(d->mg_nu1)) return -2;
    HYD_CHECK(hyd::launch_mg_fine_residual(d->g, c, d->e_r, d->e_b, d->r_r, d->r_b, d->mg_qr,
                                           d->mg_qb, d->stream));
    HYD_CHECK(hyd::launch_mg_restrict_fine(d->g, d->mg_qr, d->mg_qb, d->mg[0], d->mg_f[0],
                                           d->mg_u[0], d->stream));
    if (mg_coarse_cycle(d, 0)) return -2;
    HYD_CHECK(hyd::launch_mg_prolong_fine(d->g, d->mg[0], d->mg_u[0], d->e_r, d->e_b, d->stream));
    return sweep(d->mg_nu2);
}

// Defect correction with multigrid inner solves: the outer loop is the one
// of sor_fp32_dc (FP64 residual, phi += e, relative-residual stop), the
// inner solve a number of V-cycles sized from the previous round's
// contraction to aim at a 1e-4 reduction.  last_iterations counts V-cycles.
int dc_multigrid(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    if (ensure_dc_buffers(d) || ensure_mg(s)) return -2;
    const SorCoef c = coefficients(s);
    double rel;
    if (start_residual(s, &rel)) return -2;
    int cycles = 0, k = 3;
    const double tol = s->tolerance;
    for (int outer = 0; outer < 64 && rel > tol && cycles < s->max_iterations; ++outer) {
        const int nk = std::min(k, s->max_iterations - cycles);
        for (int q = 0; q < nk; ++q)
            if (mg_vcycle(s, c)) return -2;
        cycles += nk;
        HYD_CHECK(hyd::launch_correct(d->g, d->phi_r, d->phi_b, d->e_r, d->e_b, d->stream));
        double rnew;
        if (rel_residual_dev(s, d->r_r, d->r_b, &rnew)) return -2;
        const double red = rnew / rel;
        if (red > 0.0 && red < 1.0) {
            const double rate = pow(red, 1.0 / nk);
            const double want = std::max(1e-4, 0.5 * tol / rnew);
            k = (int)ceil(log(want) / log(rate));
            k = std::max(1, std::min(k, 16));
        } else {
            k = std::min(2 * k, 16);
        }
        rel = rnew;
    }
    s->last_iterations = cycles;
    s->last_residual = rel;
    s->last_rel_residual = rel;
    return rel <= tol ? 0 : -1;
}

// Starting potential of POISSON_NN_FP32_DC.  Candidates, in this order: the
// warm start (the potential left by the previous solve), the warm start
// plus the start-correction network's correction e of it (computed from its
// residual, and only while that residual is above the tolerance), and the
// potential-guess network's guess.  Each is scored by its relative residual
// with this density's boundary values; the smallest wins, a tie keeps the
// earlier candidate.  The potential guess costs a network pass on the full
// density at up to 128^3; it is skipped while the best start so far has
// under a quarter of the residual the guess had when last tried (its
// residual comes from grid-scale error and hardly changes from step to step:
// 5.8 on warm and 5.9 on cold 256^3 starts of the explosion, against 0.05-0.5
// for warm starts).  The corrected start is kept as phi and e: e is the
// FP32 unknown of the defect correction's inner solve, which starts from it
// (and phi += e follows that solve), while r stays phi's residual.
int choose_guess(GravitySolver3D* s, const double* d_rho) {
    GravDevice* d = s->dev;
    s->last_guess = -1;
    s->last_guess_rel_residual = -1.0;
    s->last_warm_rel_residual = -1.0;
    s->last_corr_rel_residual = -1.0;
    const bool cubic = s->nx == s->ny && s->ny == s->nz && s->dx == s->dy && s->dy == s->dz;
    if ((!s->nn_model && !s->nn_correction) || !cubic || s->bc_type != GRAV_BC_ISOLATED) return 0;
    if (ensure_dc_buffers(d)) return -2;
    double best;
    if (apply_boundary(s) || rel_residual_dev(s, d->r_r, d->r_b, &best)) return -2;
    const double ff = d->acc_host[1];  // sum of f^2 over the interior
    s->last_warm_rel_residual = best;
    bool any = false;
    bool ready = true;       // r_r / r_b hold the residual of the start in phi
    bool e_start = false;    // e holds the correction of the chosen start
    const size_t ebytes = sizeof(float) * (size_t)d->nsplit;
    if (s->nn_correction && best > s->tolerance) {
        const int rc = nn_correction_device(s->nn_correction, d->r_r, d->r_b, s->nx,
                                            s->xmax - s->xmin, d->e_r, d->e_b);
        if (rc < -1) return -2;
        if (rc == 0) {
            any = true;
            HYD_CHECK(hyd::launch_start_residual_sq(d->g, coefficients(s), d->e_r, d->e_b, d->r_r,
                                                    d->r_b, d->res_part, d->acc, d->stream));
            HYD_CHECK(hipMemcpyAsync(d->acc_host, d->acc, sizeof(double), hipMemcpyDeviceToHost,
                                     d->stream));
            HYD_CHECK(hipStreamSynchronize(d->stream));
            const double rc2 = ff > 0.0 ? sqrt(d->acc_host[0] / ff) : sqrt(d->acc_host[0]);
            s->last_corr_rel_residual = rc2;
            if (rc2 < best) {
                best = rc2;
                s->last_guess = 2;
                e_start = true;
            } else {
                HYD_CHECK(hipMemsetAsync(d->e_r, 0, ebytes, d->stream));
                HYD_CHECK(hipMemsetAsync(d->e_b, 0, ebytes, d->stream));
            }
        }
    }
    if (s->nn_model && d->guess_last_rel > 0.0 && 4.0 * best < d->guess_last_rel) {
        any = true;  // the guess cannot be expected to win: the best start stays
    } else if (s->nn_model) {
        if (!d->guess_r && (dev_alloc(&d->guess_r, d->nsplit) || dev_alloc(&d->guess_b, d->nsplit)))
            return -2;
        if (!d->nat_tmp && dev_alloc(&d->nat_tmp, d->n)) return -2;
        const double box[4] = {s->xmin, s->ymin, s->zmin, s->xmax - s->xmin};
        const int rc = nn_guess_device(s->nn_model, d_rho, s->nx, d->mom, s->G, box, d->nat_tmp);
        if (rc < -1) return -2;
        if (rc == 0) {
            any = true;
            HYD_CHECK(hyd::launch_natural_to_split(d->g, d->nat_tmp, 1.0, d->guess_r, d->guess_b,
                                                   d->stream));
            std::swap(d->phi_r, d->guess_r);
            std::swap(d->phi_b, d->guess_b);
            double rn;
            if (apply_boundary(s) || rel_residual_dev(s, nullptr, nullptr, &rn)) return -2;
            s->last_guess_rel_residual = rn;
            d->guess_last_rel = rn;
            if (rn < best) {
                best = rn;
                s->last_guess = 1;
                ready = false;
                if (e_start) {  // the corrected start lost
                    HYD_CHECK(hipMemsetAsync(d->e_r, 0, ebytes, d->stream));
                    HYD_CHECK(hipMemsetAsync(d->e_b, 0, ebytes, d->stream));
                    e_start = false;
                }
            } else {
                std::swap(d->phi_r, d->guess_r);
                std::swap(d->phi_b, d->guess_b);
            }
        }
    }
    if (any && s->last_guess < 0) s->last_guess = 0;
    if (e_start && (best <= s->tolerance || s->max_iterations < 1)) {
        // no inner solve will run: phi += e now, and the solve checks phi
        HYD_CHECK(hyd::launch_correct(d->g, d->phi_r, d->phi_b, d->e_r, d->e_b, d->stream));
        ready = false;
    }
    d->start_residual_ready = ready;
    d->start_rel = best;
    return 0;
}

int potential_device(GravitySolver3D* s, const double* d_rho) {
    GravDevice* d = s->dev;
    d->start_residual_ready = false;
    HYD_CHECK(hipEventRecord(d->ev0, d->stream));
    const double four_pi_G = 4.0 * kPi * s->G;  // gravity_solver_3d.c:527
    if (d->two_d) {
        if (s->poisson_mode != POISSON_FP64_SOR || s->criterion != SOR_CRIT_MAX_UPDATE) {
            fprintf(stderr, "Error: nz == 1 is solved by FP64 SOR with the max-update rule only.\n");
            return -2;
        }
        d->rho_2d = d_rho;
    } else {
        HYD_CHECK(hyd::launch_natural_to_split(d->g, d_rho, four_pi_G, d->rhs_r, d->rhs_b,
                                               d->stream));
    }
    if (s->bc_type == GRAV_BC_ISOLATED && moments_dev(s, d_rho)) return -2;
    int st;
    if (s->poisson_mode == POISSON_FP64_SOR) {
        st = sor_fp64(s);
    } else {
        if (s->poisson_mode == POISSON_NN_FP32_DC && choose_guess(s, d_rho)) return -2;
        st = s->dc_inner == DC_INNER_MULTIGRID ? dc_multigrid(s) : sor_fp32_dc(s);
    }
    HYD_CHECK(hipEventRecord(d->ev1, d->stream));
    HYD_CHECK(hipEventSynchronize(d->ev1));
    float ms = 0.0f;
    HYD_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    s->last_solve_ms = ms;
    s->last_status = st;
    if (st == -1) {  // gravity_solver_3d.c:733-735
        fprintf(stderr, "Warning: Gravity solver did not converge in %d iterations.\n",
                s->last_iterations);
        fprintf(stderr, "         Final residual: %.3e (tolerance: %.3e)\n", s->last_residual,
                s->tolerance);
    }
    return st;
}

int accel_device(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    if (d->two_d)
        HYD_CHECK(hyd::launch_accel2d(d->g, d->phi_r, d->gx, d->gy, d->gz, d->stream));
    else
        HYD_CHECK(hyd::launch_acceleration(d->g, d->phi_r, d->phi_b, d->gx, d->gy, d->gz, d->stream));
    d->g_valid = true;
    return 0;
}

// The potential in the natural layout (device): phi_r itself for nz == 1,
// else the colours merged into nat_tmp.
const double* natural_phi(GravDevice* d, int* err) {
    *err = 0;
    if (d->two_d) return d->phi_r;
    if (!d->nat_tmp && dev_alloc(&d->nat_tmp, d->n)) {
        *err = -2;
        return nullptr;
    }
    if (hyd::launch_split_to_natural(d->g, d->phi_r, d->phi_b, d->nat_tmp, d->stream) !=
        hipSuccess) {
        *err = -2;
        return nullptr;
    }
    return d->nat_tmp;
}

// A natural-layout potential (host) into the device potential.
int upload_phi(GravDevice* d, const double* phi) {
    const size_t b = sizeof(double) * d->n;
    if (d->two_d) {
        HYD_CHECK(hipMemcpyAsync(d->phi_r, phi, b, hipMemcpyHostToDevice, d->stream));
        return 0;
    }
    if (!d->nat_tmp && dev_alloc(&d->nat_tmp, d->n)) return -2;
    HYD_CHECK(hipMemcpyAsync(d->nat_tmp, phi, b, hipMemcpyHostToDevice, d->stream));
    HYD_CHECK(hyd::launch_natural_to_split(d->g, d->nat_tmp, 1.0, d->phi_r, d->phi_b, d->stream));
    return 0;
}

int download_phi(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    int err;
    const double* nat = natural_phi(d, &err);
    if (err) return err;
    HYD_CHECK(hipMemcpyAsync(s->phi, nat, sizeof(double) * d->n, hipMemcpyDeviceToHost,
                             d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int download_g(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    const size_t b = sizeof(double) * d->n;
    HYD_CHECK(hipMemcpyAsync(s->gx, d->gx, b, hipMemcpyDeviceToHost, d->stream));
    HYD_CHECK(hipMemcpyAsync(s->gy, d->gy, b, hipMemcpyDeviceToHost, d->stream));
    HYD_CHECK(hipMemcpyAsync(s->gz, d->gz, b, hipMemcpyDeviceToHost, d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int upload_rho(GravitySolver3D* s, const double* rho) {
    GravDevice* d = s->dev;
    if (!d->rho_tmp && dev_alloc(&d->rho_tmp, d->n)) return -2;
    HYD_CHECK(hipMemcpyAsync(d->rho_tmp, rho, sizeof(double) * d->n, hipMemcpyHostToDevice,
                             d->stream));
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" {

GravitySolver3D* gravity_solver_create(int nx, int ny, int nz, double xmin, double xmax,
                                       double ymin, double ymax, double zmin, double zmax,
                                       double G) {
    if (nx < 3 || ny < 3 || nz < 1 || nz == 2) {
        fprintf(stderr,
                "Error: gravity solver needs nx, ny >= 3 and nz == 1 or >= 3 on the device "
                "(got %d x %d x %d).\n",
                nx, ny, nz);
        return NULL;
    }
    GravitySolver3D* s = (GravitySolver3D*)calloc(1, sizeof(GravitySolver3D));
    if (!s) {
        fprintf(stderr, "Error: Failed to allocate gravity solver structure.\n");
        return NULL;
    }
    // gravity_solver_3d.c:60-86
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->xmin = xmin; s->xmax = xmax; s->ymin = ymin; s->ymax = ymax; s->zmin = zmin; s->zmax = zmax;
    s->dx = (xmax - xmin) / nx;
    s->dy = (ymax - ymin) / ny;
    s->dz = (zmax - zmin) / nz;
    s->G = G;
    s->omega = 1.5;
    s->max_iterations = 10000;
    s->tolerance = 1.0e-6;
    s->bc_type = GRAV_BC_ISOLATED;
    s->multipole_lmax = 2;
    s->poisson_mode = POISSON_FP64_SOR;
    s->criterion = SOR_CRIT_MAX_UPDATE;
    s->dc_inner = DC_INNER_SOR;
    s->moment_order = MOMENT_ORDER_AUTO;
    const long n = (long)nx * ny * nz;
    s->phi = (double*)calloc(n, sizeof(double));
    s->rho = NULL;
    s->gx = (double*)calloc(n, sizeof(double));
    s->gy = (double*)calloc(n, sizeof(double));
    s->gz = (double*)calloc(n, sizeof(double));
    if (!s->phi || !s->gx || !s->gy || !s->gz || dev_create(s) != 0) {
        fprintf(stderr, "Error: Failed to allocate gravity solver arrays.\n");
        gravity_solver_destroy(s);
        return NULL;
    }
    return s;
}

void gravity_solver_destroy(GravitySolver3D* s) {
    if (!s) return;
    if (s->dev && s->dev->owns_nn) nn_model_destroy(s->nn_model);
    if (s->dev && s->dev->owns_corr) nn_model_destroy(s->nn_correction);
    free(s->phi);
    free(s->gx);
    free(s->gy);
    free(s->gz);
    dev_free(s->dev);
    free(s);
}

int gravity_solve_potential(GravitySolver3D* s, const double* rho_field) {
    s->rho = (double*)rho_field;  // gravity_solver_3d.c:531
    if (upload_rho(s, rho_field)) return -2;
    int st = potential_device(s, s->dev->rho_tmp);
    if (st < -1) return st;
    if (download_phi(s)) return -2;
    return st;
}

void gravity_compute_acceleration(GravitySolver3D* s) {
    if (accel_device(s) == 0) (void)download_g(s);
}

void gravity_set_solver_params(GravitySolver3D* s, double omega, int max_iter, double tolerance) {
    s->omega = omega;
    s->max_iterations = max_iter;
    s->tolerance = tolerance;
}

void gravity_set_boundary_type(GravitySolver3D* s, GravBCType bc_type) { s->bc_type = bc_type; }

void gravity_get_diagnostics(const GravitySolver3D* s, int* iterations, double* residual) {
    if (iterations) *iterations = s->last_iterations;
    if (residual) *residual = s->last_residual;
}

int gravity_solve(GravitySolver3D* s, const double* rho_field) {
    s->rho = (double*)rho_field;
    if (upload_rho(s, rho_field)) return -2;
    int st = potential_device(s, s->dev->rho_tmp);
    if (st == 0 && accel_device(s)) return -2;  // gravity_solver_3d.c:905-910
    if (st < -1) return st;
    if (download_phi(s) || download_g(s)) return -2;
    return st;
}

void gravity_write_potential(const GravitySolver3D* s, const char* filename) {
    // gravity_solver_3d.c:913-943, from the host arrays
    FILE* fp = fopen(filename, "w");
    if (!fp) {
        fprintf(stderr, "Error: Could not open file %s for writing potential.\n", filename);
        return;
    }
    fprintf(fp, "x,y,z,phi,rho,gx,gy,gz\n");
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) {
                const long idx = IDX3D_GRAV(i, j, k, (long)s->ny, s->nz);
                const double x = s->xmin + (i + 0.5) * s->dx;
                const double y = s->ymin + (j + 0.5) * s->dy;
                const double z = s->zmin + (k + 0.5) * s->dz;
                const double r = s->rho ? s->rho[idx] : 0.0;
                fprintf(fp, "%.6f,%.6f,%.6f,%.6e,%.6e,%.6e,%.6e,%.6e\n", x, y, z, s->phi[idx], r,
                        s->gx[idx], s->gy[idx], s->gz[idx]);
            }
    fclose(fp);
    printf("Wrote gravitational potential to %s\n", filename);
}

// ------------------------------------------------------------- extensions --
int gravity_set_multipole_lmax(GravitySolver3D* s, int lmax) {
    if (lmax < 2 || lmax > hyd::MP_LMAX_MAX) {
        fprintf(stderr, "gravity_set_multipole_lmax: lmax must be in [2, %d]\n", hyd::MP_LMAX_MAX);
        return -1;
    }
    s->multipole_lmax = lmax;
    return 0;
}

int gravity_set_poisson_mode(GravitySolver3D* s, int mode) {
    if (mode != POISSON_FP64_SOR && mode != POISSON_FP32_DC && mode != POISSON_NN_FP32_DC)
        return -1;
    if (s->dev->two_d && mode != POISSON_FP64_SOR) {
        fprintf(stderr, "gravity_set_poisson_mode: nz == 1 is solved by FP64 SOR only\n");
        return -1;
    }
    s->poisson_mode = mode;
    return 0;
}

int gravity_set_moment_order(GravitySolver3D* s, int order) {
    if (order != MOMENT_ORDER_TREE && order != MOMENT_ORDER_SEQUENTIAL &&
        order != MOMENT_ORDER_AUTO)
        return -1;
    s->moment_order = order;
    return 0;
}

int gravity_set_dc_inner(GravitySolver3D* s, int inner) {
    if (inner != DC_INNER_SOR && inner != DC_INNER_MULTIGRID) return -1;
    if (inner == DC_INNER_MULTIGRID && !mg_supported(s)) {
        fprintf(stderr, "gravity_set_dc_inner: multigrid needs even nx, ny, nz >= 8 (got %d x %d x %d)\n",
                s->nx, s->ny, s->nz);
        return -1;
    }
    s->dc_inner = inner;
    return 0;
}

int gravity_set_nn_model(GravitySolver3D* s, void* model) {
    if (model && nn_model_kind(model, nullptr) != 0) {
        fprintf(stderr, "gravity_set_nn_model: the network is not a potential guess\n");
        return -1;
    }
    if (s->dev->owns_nn && s->nn_model != model) {
        nn_model_destroy(s->nn_model);
        s->dev->owns_nn = false;
    }
    s->nn_model = model;
    s->dev->guess_last_rel = -1.0;  // the next solve tries the new network
    return 0;
}

int gravity_set_nn_correction(GravitySolver3D* s, void* model) {
    if (model && nn_model_kind(model, nullptr) != 1) {
        fprintf(stderr, "gravity_set_nn_correction: the network is not a start correction\n");
        return -1;
    }
    if (s->dev->owns_corr && s->nn_correction != model) {
        nn_model_destroy(s->nn_correction);
        s->dev->owns_corr = false;
    }
    s->nn_correction = model;
    return 0;
}

int gravity_set_criterion(GravitySolver3D* s, int criterion) {
    if (criterion != SOR_CRIT_MAX_UPDATE && criterion != SOR_CRIT_REL_RESID) return -1;
    if (s->dev->two_d && criterion != SOR_CRIT_MAX_UPDATE) {
        fprintf(stderr, "gravity_set_criterion: nz == 1 stops by the max-update rule only\n");
        return -1;
    }
    s->criterion = criterion;
    return 0;
}

int gravity_sync_to_device(GravitySolver3D* s) {
    GravDevice* d = s->dev;
    if (upload_phi(d, s->phi)) return -2;
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int gravity_sync_to_host(GravitySolver3D* s) {
    if (download_phi(s)) return -2;
    return download_g(s);
}

int gravity_moments(GravitySolver3D* s, const double* rho, double* out13, double* out_hi) {
    GravDevice* d = s->dev;
    if (upload_rho(s, rho)) return -2;
    if (moments_dev(s, d->rho_tmp)) return -2;
    const int nhi = hyd::mp_n_hi(s->multipole_lmax);
    double buf[13 + 128];
    HYD_CHECK(hipMemcpyAsync(buf, d->mom, sizeof(double) * (13 + nhi), hipMemcpyDeviceToHost,
                             d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    memcpy(out13, buf, 13 * sizeof(double));
    if (out_hi) memcpy(out_hi, buf + 13, nhi * sizeof(double));
    return 0;
}

int gravity_boundary_potential(GravitySolver3D* s, const double* rho, double* phi_out) {
    GravDevice* d = s->dev;
    if (upload_rho(s, rho)) return -2;
    if (s->bc_type == GRAV_BC_ISOLATED && moments_dev(s, d->rho_tmp)) return -2;
    if (apply_boundary(s)) return -2;
    int err;
    const double* nat = natural_phi(d, &err);
    if (err) return err;
    HYD_CHECK(hipMemcpyAsync(phi_out, nat, sizeof(double) * d->n, hipMemcpyDeviceToHost,
                             d->stream));
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

int gravity_rel_residual(GravitySolver3D* s, const double* rho, double* rel) {
    GravDevice* d = s->dev;
    if (d->two_d) {
        fprintf(stderr, "gravity_rel_residual: not provided for nz == 1\n");
        return -1;
    }
    if (upload_rho(s, rho)) return -2;
    HYD_CHECK(hyd::launch_natural_to_split(d->g, d->rho_tmp, 4.0 * kPi * s->G, d->rhs_r, d->rhs_b,
                                           d->stream));
    return rel_residual_dev(s, nullptr, nullptr, rel);
}

int gravity_solve_device(GravitySolver3D* s, const double* d_rho) {
    return grav_internal_solve(s, d_rho, s->dev->stream);
}

}  // extern "C"

// Hydro-side entry: potential and, when it converged, acceleration, all on
// the device (solve_gravity_field, hydro_solver_3d.c:558-591; a solve that
// did not converge leaves g as it was, gravity_solver_3d.c:907).
int grav_internal_solve(GravitySolver3D* s, const double* d_rho, hipStream_t) {
    int st = potential_device(s, d_rho);
    if (st == 0 && accel_device(s)) return -2;
    return st;
}

int grav_internal_restore(GravitySolver3D* s, const double* phi, const double* gx,
                          const double* gy, const double* gz) {
    GravDevice* d = s->dev;
    const size_t b = sizeof(double) * d->n;
    if (phi && upload_phi(d, phi)) return -2;
    if (gx && gy && gz) {
        HYD_CHECK(hipMemcpyAsync(d->gx, gx, b, hipMemcpyHostToDevice, d->stream));
        HYD_CHECK(hipMemcpyAsync(d->gy, gy, b, hipMemcpyHostToDevice, d->stream));
        HYD_CHECK(hipMemcpyAsync(d->gz, gz, b, hipMemcpyHostToDevice, d->stream));
        d->g_valid = true;
    }
    HYD_CHECK(hipStreamSynchronize(d->stream));
    return 0;
}

const double* grav_internal_g(GravitySolver3D* s, int comp) {
    GravDevice* d = s->dev;
    return comp == 0 ? d->gx : (comp == 1 ? d->gy : d->gz);
}
