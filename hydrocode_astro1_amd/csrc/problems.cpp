// problems.cpp -- the reference's initial-condition problems, on the host.
//
// set_initial_conditions_3d (hydro_solver_3d.c:1383-1515) and
// set_initial_conditions_3d_from_deck (hydro_solver_3d.c:2092-2133) fill the
// state on the host before the first upload, as the reference does.  The
// generators call libm (tanh, exp, sin, cos, sqrt, log, acos, atan2) and a
// 32-bit linear congruential generator, so every formula below keeps the
// reference's operand order: with the same libm the initial state is the
// reference's bit for bit (tests/test_problems.py compares them against the
// compiled reference).  The device plays no part; running these on the GPU
// would change the libm results and gain nothing for a one-off set-up.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hydro_mi355x.h"

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI
constexpr double kRhoFloor = RHO_FLOOR;
constexpr double kPFloor = P_FLOOR;

inline double maxd(double a, double b) { return a > b ? a : b; }  // the reference's MAX

struct Prim {
    double rho, u, v, w, p;
};

// Grid geometry and EOS of the state being filled.
struct Box {
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax, dx, dy, dz, gamma, G;
    State3D* U;

    long at(int i, int j, int k) const { return ((long)i * ny + j) * nz + k; }
    double xc(int i) const { return xmin + (i + 0.5) * dx; }
    double yc(int j) const { return ymin + (j + 0.5) * dy; }
    double zc(int k) const { return zmin + (k + 0.5) * dz; }

    // prim_to_cons_3d, hydro_solver_3d.c:177-189
    void put(long c, const Prim& P) const {
        State3D& s = U[c];
        s.rho = P.rho;
        s.rhou = P.rho * P.u;
        s.rhov = P.rho * P.v;
        s.rhow = P.rho * P.w;
        const double ke = 0.5 * P.rho * (P.u * P.u + P.v * P.v + P.w * P.w);
        s.E = P.p / (gamma - 1.0) + ke;
    }
    // cons_to_prim_3d, hydro_solver_3d.c:162-175 (with its floors)
    Prim get(long c) const {
        const State3D& s = U[c];
        Prim P;
        P.rho = maxd(s.rho, kRhoFloor);
        P.u = s.rhou / P.rho;
        P.v = s.rhov / P.rho;
        P.w = s.rhow / P.rho;
        const double ke = 0.5 * P.rho * (P.u * P.u + P.v * P.v + P.w * P.w);
        P.p = (gamma - 1.0) * (s.E - ke);
        P.p = maxd(P.p, kPFloor);
        return P;
    }
    // sound_speed_3d, hydro_solver_3d.c:191-193
    double cs(const Prim& P) const { return std::sqrt(gamma * P.p / P.rho); }

    // Visit every cell in the reference's loop order (i, then j, then k).
    template <class F>
    void each(F f) const {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) f(at(i, j, k), xc(i), yc(j), zc(k));
    }
};

// A radial profile file: '#' lines skipped, then "r rho u" per line
// (read_stellar_profile, hydro_solver_3d.c:825-878).
struct Profile {
    std::vector<double> r, rho, u;
    double R = 0.0;  // first radius where rho < 1e-4 rho[0], else the last radius

    bool load(const char* path) {
        FILE* fp = std::fopen(path, "r");
        if (!fp) {
            std::fprintf(stderr, "Error: Could not open stellar profile file '%s'\n", path);
            return false;
        }
        char line[512];
        while (std::fgets(line, sizeof line, fp)) {
            if (line[0] == '#' || std::strlen(line) <= 5) continue;
            double a = 0.0, b = 0.0, c = 0.0;
            std::sscanf(line, "%lf %lf %lf", &a, &b, &c);
            r.push_back(a);
            rho.push_back(b);
            u.push_back(c);
        }
        std::fclose(fp);
        if (r.empty()) return false;
        R = r.back();
        const double cut = 1e-4 * rho[0];
        for (size_t j = 0; j < r.size(); ++j)
            if (rho[j] < cut) {
                R = r[j];
                break;
            }
        std::printf("Read stellar profile %s: %zu points, R = %.4f, rho_c = %.4e, u_c = %.4e\n",
                    path, r.size(), R, rho[0], u[0]);
        return true;
    }

    // Linear in r between the bracketing points; the centre value inside
    // r[0]; 1e-6 for both beyond the last point (hydro_solver_3d.c:888-923).
    void at(double x, double& rho_o, double& u_o) const {
        const size_t n = r.size();
        if (x <= r[0]) {
            rho_o = rho[0];
            u_o = u[0];
            return;
        }
        if (x >= r[n - 1]) {
            rho_o = 1e-6;
            u_o = 1e-6;
            return;
        }
        size_t lo = 0;
        for (size_t i = 0; i + 1 < n; ++i)
            if (x >= r[i] && x < r[i + 1]) {
                lo = i;
                break;
            }
        const double f = (x - r[lo]) / (r[lo + 1] - r[lo]);
        rho_o = rho[lo] + f * (rho[lo + 1] - rho[lo]);
        u_o = u[lo] + f * (u[lo + 1] - u[lo]);
    }

    // 4 pi rho_mid r_mid^2 dr summed over intervals (hydro_solver_3d.c:1048-1058)
    double mass() const {
        double M = 0.0;
        for (size_t i = 0; i + 1 < r.size(); ++i) {
            const double dr = r[i + 1] - r[i];
            const double rm = 0.5 * (r[i] + r[i + 1]);
            const double dm = 0.5 * (rho[i] + rho[i + 1]);
            M += 4.0 * kPi * dm * rm * rm * dr;
        }
        return M;
    }
};

inline double dist(double x, double y, double z, double a, double b, double c) {
    // sqrt(pow(x - a, 2) + ...): pow(t, 2) is t * t exactly
    const double p = x - a, q = y - b, s = z - c;
    return std::sqrt(p * p + q * q + s * s);
}

// ---------------------------------------------------------------- problems --

// hydro_solver_3d.c:1384-1442
void kelvin_helmholtz(const Box& g) {
    const double rho1 = 2.0, rho2 = 1.0, v1 = -0.5, v2 = 0.5, p0 = 2.5, w0 = 0.1, A = 0.1;
    const double kx = 2.0 * kPi / 2.0, kz = 2.0 * kPi / 2.0;
    g.each([&](long c, double x, double y, double z) {
        const double f = 0.5 * (1.0 + std::tanh((y - 0.0) / w0));
        const double env = std::exp(-y * y / (4.0 * w0 * w0));
        Prim P;
        P.rho = rho1 * f + rho2 * (1.0 - f);
        P.u = v1 * f + v2 * (1.0 - f);
        P.v = A * std::sin(kx * x) * std::cos(kz * z) * env;
        P.w = 0.5 * A * std::cos(kx * x) * std::sin(kz * z) * env;
        P.p = p0;
        g.put(c, P);
    });
}

// hydro_solver_3d.c:1444-1490
void blast_wave(const Box& g) {
    const double cx = 0.5 * (g.xmin + g.xmax), cy = 0.5 * (g.ymin + g.ymax),
                 cz = 0.5 * (g.zmin + g.zmax);
    const double rb = 0.1;
    const double V = (4.0 / 3.0) * kPi * rb * rb * rb;
    const double pb = (g.gamma - 1.0) * 1.0 / V;
    g.each([&](long c, double x, double y, double z) {
        const double r = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
        g.put(c, Prim{1.0, 0.0, 0.0, 0.0, r < rb ? pb : 1.0e-5});
    });
}

// hydro_solver_3d.c:927-1002: the profile centred in the box, at rest
void profile_star(const Box& g, const char* path) {
    Profile pr;
    if (!pr.load(path)) {
        std::fprintf(stderr, "Failed to read profile. Using uniform density instead.\n");
        g.each([&](long c, double, double, double) { g.put(c, Prim{1.0, 0.0, 0.0, 0.0, 1.0}); });
        return;
    }
    const double cx = 0.5 * (g.xmin + g.xmax), cy = 0.5 * (g.ymin + g.ymax),
                 cz = 0.5 * (g.zmin + g.zmax);
    g.each([&](long c, double x, double y, double z) {
        double rho, u;
        pr.at(dist(x, y, z, cx, cy, cz), rho, u);
        Prim P{rho, 0.0, 0.0, 0.0, (g.gamma - 1.0) * rho * u};
        P.rho = maxd(P.rho, kRhoFloor);
        P.p = maxd(P.p, kPFloor);
        g.put(c, P);
    });
}

struct Star {
    const Profile* pr;
    double x, y, z, vx, vy, vz;
};

// The two-star map shared by both binary set-ups (hydro_solver_3d.c:1134-1206
// and 1300-1372): each cell takes the star whose radius it lies inside, a
// density-weighted blend where both overlap, the floors elsewhere.
void two_stars(const Box& g, const Star& a, const Star& b) {
    g.each([&](long c, double x, double y, double z) {
        const double r1 = dist(x, y, z, a.x, a.y, a.z), r2 = dist(x, y, z, b.x, b.y, b.z);
        double rho1, u1, rho2, u2;
        a.pr->at(r1, rho1, u1);
        b.pr->at(r2, rho2, u2);
        const bool in1 = r1 < a.pr->R, in2 = r2 < b.pr->R;
        Prim P;
        if (in1 && !in2) {
            P = Prim{rho1, a.vx, a.vy, a.vz, (g.gamma - 1.0) * rho1 * u1};
        } else if (in2 && !in1) {
            P = Prim{rho2, b.vx, b.vy, b.vz, (g.gamma - 1.0) * rho2 * u2};
        } else if (in1 && in2) {
            const double tot = rho1 + rho2;
            const double w1 = rho1 / tot, w2 = rho2 / tot;
            const double p1 = (g.gamma - 1.0) * rho1 * u1, p2 = (g.gamma - 1.0) * rho2 * u2;
            P = Prim{tot, w1 * a.vx + w2 * b.vx, w1 * a.vy + w2 * b.vy, w1 * a.vz + w2 * b.vz,
                     p1 + p2};
        } else {
            P = Prim{kRhoFloor, 0.0, 0.0, 0.0, kPFloor};
        }
        P.rho = maxd(P.rho, kRhoFloor);
        P.p = maxd(P.p, kPFloor);
        g.put(c, P);
    });
}

// One profile for both stars, 3 R apart on the x axis, equal masses on a
// circular orbit about the origin (hydro_solver_3d.c:1218-1378).
void binary_same(const Box& g, const char* path) {
    Profile pr;
    if (!pr.load(path)) {
        std::fprintf(stderr, "Failed to read profile. Cannot create binary system.\n");
        return;
    }
    const double M = pr.mass();
    const double sep = 3.0 * pr.R;
    const double v_orb = std::sqrt(g.G * (2.0 * M) / sep);
    const double v = v_orb * 0.5;
    std::printf("Binary: M = %.4e, separation %.4f, v_orb = %.4e, period %.4e\n", M, sep, v_orb,
                2.0 * kPi * sep / v_orb);
    two_stars(g, Star{&pr, -sep / 2.0, 0.0, 0.0, 0.0, +v, 0.0},
              Star{&pr, sep / 2.0, 0.0, 0.0, 0.0, -v, 0.0});
}

// Two profile files; positions and velocities from the deck, or a circular
// orbit about the centre of mass along x (hydro_solver_3d.c:1008-1210).
void binary_dual(const Box& g, const InputDeck& d) {
    Profile p1, p2;
    const bool ok1 = p1.load(d.star1_profile), ok2 = p2.load(d.star2_profile);
    if (!ok1 || !ok2) {
        std::fprintf(stderr, "Failed to read one or both profiles. Cannot create binary system.\n");
        return;
    }
    const double M1 = p1.mass(), M2 = p2.mass();
    Star a{&p1, d.star1_x, d.star1_y, d.star1_z, d.star1_vx, d.star1_vy, d.star1_vz};
    Star b{&p2, d.star2_x, d.star2_y, d.star2_z, d.star2_vx, d.star2_vy, d.star2_vz};
    if (d.binary_auto_orbit) {
        const double Mt = M1 + M2, sep = d.binary_separation;
        const double v_orb = std::sqrt(g.G * Mt / sep);
        a = Star{&p1, -(sep * M2 / Mt), 0.0, 0.0, 0.0, +(v_orb * M2 / Mt), 0.0};
        b = Star{&p2, sep * M1 / Mt, 0.0, 0.0, 0.0, -(v_orb * M1 / Mt), 0.0};
        std::printf("Binary: M1 = %.4e, M2 = %.4e, v_orb = %.4e, period %.4e\n", M1, M2, v_orb,
                    2.0 * kPi * sep / v_orb);
    }
    two_stars(g, a, b);
}

// The profile in the octant x, y, z >= 0 with an explosion inside
// explosion_radius: thermal (energy / cell count added per cell as internal
// energy) or kinetic (a radial speed from energy and enclosed mass)
// (hydro_solver_3d.c:1864-2020).
void octant_explosion(const Box& g, const InputDeck& d) {
    Profile pr;
    if (!pr.load(d.stellar_profile)) {
        std::fprintf(stderr, "Failed to read stellar profile. Cannot create stellar octant.\n");
        return;
    }
    const double Re = d.explosion_radius;
    int count = 0;
    g.each([&](long, double x, double y, double z) {
        if (x < 0.0 || y < 0.0 || z < 0.0) return;
        if (std::sqrt(x * x + y * y + z * z) < Re) ++count;
    });
    const bool thermal = std::strcmp(d.explosion_type, "thermal") == 0;
    double e_cell = 0.0, v_kick = 0.0;
    if (thermal) {
        e_cell = d.explosion_energy / (double)count;
    } else {
        double mass = 0.0;
        g.each([&](long, double x, double y, double z) {
            if (x < 0.0 || y < 0.0 || z < 0.0) return;
            const double r = std::sqrt(x * x + y * y + z * z);
            if (r < Re) {
                double rho, u;
                pr.at(r, rho, u);
                mass += rho * g.dx * g.dy * g.dz;
            }
        });
        v_kick = std::sqrt(2.0 * d.explosion_energy / mass);
    }
    std::printf("Octant explosion: %d cells within r < %.3f, %s, %s %.3e\n", count, Re,
                thermal ? "thermal" : "kinetic", thermal ? "energy per cell" : "radial speed",
                thermal ? e_cell : v_kick);
    g.each([&](long c, double x, double y, double z) {
        Prim P;
        if (x >= 0.0 && y >= 0.0 && z >= 0.0) {
            const double r = std::sqrt(x * x + y * y + z * z);
            double rho, u;
            pr.at(r, rho, u);
            P = Prim{rho, 0.0, 0.0, 0.0, rho * u * (g.gamma - 1.0)};
            if (r < Re) {
                if (thermal) {
                    double ui = u;
                    ui += e_cell / (rho * g.dx * g.dy * g.dz);
                    P.p = rho * ui * (g.gamma - 1.0);
                } else if (r > 1e-10) {
                    P.u = v_kick * (x / r);
                    P.v = v_kick * (y / r);
                    P.w = v_kick * (z / r);
                }
            }
        } else {
            P = Prim{kRhoFloor, 0.0, 0.0, 0.0, kPFloor};
        }
        g.put(c, P);
    });
}

// The reference's generator (hydro_solver_3d.c:1522-1539): a 32-bit linear
// congruential sequence (Numerical Recipes constants) and Box-Muller.
struct Lcg {
    unsigned long s;
    double uniform() {
        s = (s * 1664525UL + 1013904223UL) & 0xFFFFFFFFUL;
        return (double)s / 4294967296.0;
    }
    double normal() {
        const double a = uniform();
        const double b = uniform();
        return std::sqrt(-2.0 * std::log(a + 1e-10)) * std::cos(2.0 * kPi * b);
    }
};

struct Wave {
    double kx, ky, kz, ax, ay, az, phase;
};

// A wave vector of random length in [kmin, kmax] and direction uniform on the
// sphere (hydro_solver_3d.c:1614-1623, 1700-1707).
void random_direction(Lcg& g, double kmin, double kmax, Wave& w) {
    const double k = kmin + (kmax - kmin) * g.uniform();
    const double ph = 2.0 * kPi * g.uniform();
    const double ct = 2.0 * g.uniform() - 1.0;
    const double st = std::sqrt(1.0 - ct * ct);
    w.kx = k * st * std::cos(ph);
    w.ky = k * st * std::sin(ph);
    w.kz = k * ct;
}

inline bool outside_octant(double x, double y, double z) { return x < 0.0 || y < 0.0 || z < 0.0; }

// Velocity += amplitude c_s (sum_n a_n cos(k_n.x + phase_n)) / sqrt(N) per
// component (hydro_solver_3d.c:1596-1680).
void perturb_velocity(const Box& g, Lcg& rng, double amp, double kmin, double kmax, int n) {
    std::vector<Wave> m(n);
    for (auto& w : m) {
        random_direction(rng, kmin, kmax, w);
        w.ax = rng.normal();
        w.ay = rng.normal();
        w.az = rng.normal();
        w.phase = 2.0 * kPi * rng.uniform();
    }
    g.each([&](long c, double x, double y, double z) {
        if (outside_octant(x, y, z)) return;
        Prim P = g.get(c);
        const double cs = g.cs(P);
        double du = 0.0, dv = 0.0, dw = 0.0;
        for (const auto& w : m) {
            const double wave = std::cos(w.kx * x + w.ky * y + w.kz * z + w.phase);
            du += w.ax * wave;
            dv += w.ay * wave;
            dw += w.az * wave;
        }
        const double norm = std::sqrt((double)n);
        P.u += (du / norm) * amp * cs;
        P.v += (dv / norm) * amp * cs;
        P.w += (dw / norm) * amp * cs;
        g.put(c, P);
    });
}

// rho *= max(1 + amplitude (sum_n a_n cos(...)) / sqrt(N), 0.1)
// (hydro_solver_3d.c:1683-1753).
void perturb_density(const Box& g, Lcg& rng, double amp, double kmin, double kmax, int n) {
    std::vector<Wave> m(n);
    for (auto& w : m) {
        random_direction(rng, kmin, kmax, w);
        w.ax = rng.normal();
        w.phase = 2.0 * kPi * rng.uniform();
    }
    g.each([&](long c, double x, double y, double z) {
        if (outside_octant(x, y, z)) return;
        Prim P = g.get(c);
        double s = 0.0;
        for (const auto& w : m) s += w.ax * std::cos(w.kx * x + w.ky * y + w.kz * z + w.phase);
        const double norm = std::sqrt((double)n);
        double f = 1.0 + (s / norm) * amp;
        if (f < 0.1) f = 0.1;
        P.rho *= f;
        g.put(c, P);
    });
}

// Associated Legendre function by the upward recurrence in l, with the
// Condon-Shortley sign (hydro_solver_3d.c:1543-1573).
double legendre(int l, int m, double x) {
    if (m < 0 || m > l) return 0.0;
    if (x > 1.0) x = 1.0;
    if (x < -1.0) x = -1.0;
    double pmm = 1.0;
    if (m > 0) {
        const double s = std::sqrt((1.0 - x) * (1.0 + x));
        double f = 1.0;
        for (int i = 1; i <= m; ++i) {
            pmm *= -f * s;
            f += 2.0;
        }
    }
    if (l == m) return pmm;
    double p1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) return p1;
    double pl = 0.0;
    for (int ll = m + 2; ll <= l; ++ll) {
        pl = ((2 * ll - 1) * x * p1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = p1;
        p1 = pl;
    }
    return pl;
}

// sqrt((2l+1)/4pi) P_l^|m|(cos theta) times cos(m phi), sin(|m| phi) or 1
// (hydro_solver_3d.c:1576-1593).
double real_harmonic(int l, int m, double theta, double phi) {
    const double P = legendre(l, m < 0 ? -m : m, std::cos(theta));
    const double norm = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi));
    if (m > 0) return norm * P * std::cos(m * phi);
    if (m < 0) return norm * P * std::sin((-m) * phi);
    return norm * P;
}

// rho *= 1 + a_rho Y; a radial velocity a_v c_s Y (hydro_solver_3d.c:1756-1814).
void perturb_harmonic(const Box& g, int l, int m, double a_rho, double a_vel) {
    g.each([&](long c, double x, double y, double z) {
        if (outside_octant(x, y, z)) return;
        const double r = std::sqrt(x * x + y * y + z * z);
        if (r < 1e-10) return;
        const double Y = real_harmonic(l, m, std::acos(z / r), std::atan2(y, x));
        Prim P = g.get(c);
        const double cs = g.cs(P);
        P.rho *= (1.0 + a_rho * Y);
        const double vr = a_vel * cs * Y;
        P.u += vr * (x / r);
        P.v += vr * (y / r);
        P.w += vr * (z / r);
        g.put(c, P);
    });
}

// hydro_solver_3d.c:1817-1861: seeded once, velocity, then density, then the
// harmonic mode, each when enabled.
void perturbations(const Box& g, const InputDeck& d) {
    Lcg rng{(unsigned long)d.perturbation_seed};
    if (d.perturb_velocity_amplitude > 0.0 && d.perturb_velocity_modes > 0)
        perturb_velocity(g, rng, d.perturb_velocity_amplitude, d.perturb_velocity_kmin,
                         d.perturb_velocity_kmax, d.perturb_velocity_modes);
    if (d.perturb_density_amplitude > 0.0 && d.perturb_density_modes > 0)
        perturb_density(g, rng, d.perturb_density_amplitude, d.perturb_density_kmin,
                        d.perturb_density_kmax, d.perturb_density_modes);
    if (d.enable_sph_harmonic_mode)
        perturb_harmonic(g, d.sph_harmonic_l, d.sph_harmonic_m, d.sph_harmonic_amp_rho,
                         d.sph_harmonic_amp_vel);
    std::printf("Perturbations applied (seed %d).\n", d.perturbation_seed);
}

// A torus of given density and pressure about the z axis, rotating
// azimuthally, overwriting the octant cells it covers (hydro_solver_3d.c:2026-2088).
void torus(const Box& g, const InputDeck& d) {
    const double R = d.torus_major_radius, rt = d.torus_minor_radius, vr = d.torus_rotation_velocity;
    int cells = 0;
    g.each([&](long c, double x, double y, double z) {
        if (outside_octant(x, y, z)) return;
        const double rc = std::sqrt(x * x + y * y);
        if (std::sqrt((rc - R) * (rc - R) + z * z) > rt) return;
        Prim P{d.torus_density, 0.0, 0.0, 0.0, d.torus_pressure};
        if (rc > 1e-10) {
            P.u = -vr * y / rc;
            P.v = vr * x / rc;
        }
        g.put(c, P);
        ++cells;
    });
    std::printf("Torus added: %d cells.\n", cells);
}

}  // namespace

extern "C" int hydro_fill_initial_state(int nx, int ny, int nz, double xmin, double xmax,
                                        double ymin, double ymax, double zmin, double zmax,
                                        double gamma, double G, int problem, const InputDeck* deck,
                                        State3D* U) {
    if (nx < 1 || ny < 1 || nz < 1 || !U) return -1;
    Box g{nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax,
          (xmax - xmin) / nx, (ymax - ymin) / ny, (zmax - zmin) / nz, gamma, G, U};
    const ProblemType p = deck ? deck->problem : (ProblemType)problem;
    if (deck && p == STELLAR_OCTANT_EXPLOSION_3D) {
        octant_explosion(g, *deck);
        if (deck->enable_perturbations) perturbations(g, *deck);
        if (deck->enable_torus) torus(g, *deck);
        return 0;
    }
    if (deck && p == BINARY_STAR_3D) {
        binary_dual(g, *deck);
        return 0;
    }
    switch (p) {
        case KELVIN_HELMHOLTZ_3D: kelvin_helmholtz(g); return 0;
        case BLAST_WAVE_3D: blast_wave(g); return 0;
        case STELLAR_PROFILE_3D: profile_star(g, "stellar_profile.txt"); return 0;
        case BINARY_STAR_3D: binary_same(g, "stellar_profile.txt"); return 0;
        default: return 1;  // no generator: the state is left as it is
    }
}
