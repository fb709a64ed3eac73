// device_common.hpp -- shared device code: EOS with floors, HLLC face flux,
// double-double arithmetic and small reduction helpers.
//
// Everything here is compiled with -ffp-contract=off: each expression is
// evaluated in the written order with one rounding per operation, which is
// how the reference C solver (built by gcc -std=c11, no FMA contraction)
// evaluates it.  Division and sqrt are the IEEE correctly rounded forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HYD_RHO_FLOOR 1.0e-5  // hydro_solver_3d.h:10
#define HYD_P_FLOOR 1.0e-6    // hydro_solver_3d.h:11

namespace hyd {

// hydro_solver_3d.c:11-12: ternaries; a NaN first argument yields the second.
__device__ __forceinline__ double tmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double tmax(double a, double b) { return (a > b) ? a : b; }

// Keeps the rare IEEE recomputation a real branch: the compiler may not
// speculate across a volatile asm statement, so the common path never
// executes the full division.
#define HYD_COLD_PATH() asm volatile("")

// x / d for a divisor d that is a compile-time constant after unrolling (a
// small integer here), correctly rounded: x (1/d) for a power of two (exact),
// else one Markstein correction of q0 = RN(x RN(1/d)): the remainder x - d q0
// is exact in a fused multiply-add and RN(q0 + r RN(1/d)) is RN(x/d)
// (Markstein 1990; Muller et al., Handbook of Floating-Point Arithmetic, the
// quotient theorem), which holds while x stays clear of underflow and
// overflow; outside that range the IEEE division runs.
__device__ __forceinline__ double div_by_const(double x, double d) {
    if (d == 2.0 || d == 4.0 || d == 8.0) return x * (1.0 / d);
    const double y = 1.0 / d;
    const double q0 = x * y;
    const double r = fma(-q0, d, x);
    double q = fma(r, y, q0);
    const double ax = fabs(x);
    if (!(ax >= 0x1p-900 && ax <= 0x1p+900)) {
        HYD_COLD_PATH();
        q = x / d;
    }
    return q;
}

// Conserved state of one cell plus its primitives (P.rho itself is
// recomputed from d by the floor).
struct Cell {
    double d, mx, my, mz, e;  // rho, rho u, rho v, rho w, E
    double u, v, w, p, c;     // velocities, pressure (floored), sound speed
};

__device__ __forceinline__ double rho_floor(double d) { return tmax(d, HYD_RHO_FLOOR); }

// cons_to_prim_3d, hydro_solver_3d.c:162-175; sound_speed_3d :191-193.
__device__ __forceinline__ void prims(Cell& q, double gamma) {
    const double r = rho_floor(q.d);
    q.u = q.mx / r;
    q.v = q.my / r;
    q.w = q.mz / r;
    double ke = 0.5 * r * (q.u * q.u + q.v * q.v + q.w * q.w);
    double p = (gamma - 1.0) * (q.e - ke);
    q.p = tmax(p, HYD_P_FLOOR);
    q.c = sqrt(gamma * q.p / r);
}

// Constants of one step shared by every cell.
struct EosConst {
    double gamma, gm1, ygm1;  // gamma, gamma - 1, RN(1 / (gamma - 1))
    double yfloor;            // RN(1 / RHO_FLOOR)
};

__device__ __forceinline__ EosConst eos_const(double gamma) {
    EosConst k;
    k.gamma = gamma;
    k.gm1 = gamma - 1.0;
    k.ygm1 = 1.0 / k.gm1;
    k.yfloor = 1.0 / HYD_RHO_FLOOR;
    return k;
}

// Floors and the cons->prim->cons round trip of the update,
// hydro_solver_3d.c:765-768 with prim_to_cons_3d :177-189.
__device__ __forceinline__ void floor_roundtrip(double& d, double& mx, double& my, double& mz,
                                                double& e, const EosConst& k) {
    const double r = tmax(d, HYD_RHO_FLOOR);
    const double u = mx / r, v = my / r, w = mz / r;
    const double ke = 0.5 * r * (u * u + v * v + w * w);
    const double p = tmax(k.gm1 * (e - ke), HYD_P_FLOOR);
    const double ep = p / k.gm1;
    d = r;
    mx = r * u;
    my = r * v;
    mz = r * w;
    e = ep + ke;
}

// Reflect the velocity component normal to a face through cons->prim->cons,
// hydro_solver_3d.c:466-469 (x), :500-503 (y), :529-532 (z).
template <int DIR>
__device__ __forceinline__ void reflect(double& d, double& mx, double& my, double& mz, double& e,
                                        double gamma) {
    double r = tmax(d, HYD_RHO_FLOOR);
    double u = mx / r, v = my / r, w = mz / r;
    double ke = 0.5 * r * (u * u + v * v + w * w);
    double p = (gamma - 1.0) * (e - ke);
    p = tmax(p, HYD_P_FLOOR);
    if (DIR == 0) u = -u;
    if (DIR == 1) v = -v;
    if (DIR == 2) w = -w;
    d = r;
    mx = r * u;
    my = r * v;
    mz = r * w;
    double ke2 = 0.5 * r * (u * u + v * v + w * w);
    e = p / (gamma - 1.0) + ke2;
}

// True when one of a cell's three signal speeds |u|+c, |v|+c, |w|+c is NaN:
// the terms are >= 0 or NaN, so their sum is NaN exactly then.
__device__ __forceinline__ bool speeds_have_nan(double a, double b, double z, double c) {
    const double t = (a + b) + (z + c);
    return t != t;
}

// Largest signal speed of a cell as compute_dt_3d sees it
// (hydro_solver_3d.c:428-440): max(|u|,|v|,|w|) + c.  Adding c is monotone
// under rounding, so this equals the max of the three sums bit for bit when
// none of them is NaN; `nan` is set when one is (the reduction then takes
// the reference's order into account, nan_fold_block).
__device__ __forceinline__ double signal_speed(double d, double mx, double my, double mz,
                                               double e, double gamma, bool& nan) {
    double r = rho_floor(d);
    double u = mx / r, v = my / r, w = mz / r;
    double ke = 0.5 * r * (u * u + v * v + w * w);
    double p = tmax((gamma - 1.0) * (e - ke), HYD_P_FLOOR);
    double c = sqrt(gamma * p / r);
    double a = fabs(u), b = fabs(v), z = fabs(w);
    nan = nan || speeds_have_nan(a, b, z, c);
    double m = tmax(tmax(a, b), z);
    return m + c;
}

// The same for a state whose density is already at or above the floor.
__device__ __forceinline__ double signal_speed_floored(double r, double mx, double my, double mz,
                                                       double e, const EosConst& k, bool& nan) {
    const double u = mx / r, v = my / r, w = mz / r;
    const double ke = 0.5 * r * (u * u + v * v + w * w);
    const double p = tmax(k.gm1 * (e - ke), HYD_P_FLOOR);
    const double c = sqrt(k.gamma * p / r);
    const double a = fabs(u), b = fabs(v), z = fabs(w);
    nan = nan || speeds_have_nan(a, b, z, c);
    const double m = tmax(tmax(a, b), z);
    return m + c;
}

// The three speeds |u|+c, |v|+c, |w|+c of a cell in the reference's order
// (hydro_solver_3d.c:428-440).
__device__ __forceinline__ void cell_speeds(double d, double mx, double my, double mz, double e,
                                            double gamma, double s[3]) {
    const double r = rho_floor(d);
    const double u = mx / r, v = my / r, w = mz / r;
    const double ke = 0.5 * r * (u * u + v * v + w * w);
    const double p = tmax((gamma - 1.0) * (e - ke), HYD_P_FLOOR);
    const double c = sqrt(gamma * p / r);
    s[0] = fabs(u) + c;
    s[1] = fabs(v) + c;
    s[2] = fabs(w) + c;
}

// Velocities of the gravity source (hydro_solver_3d.c:612-614): no floor.
__device__ __forceinline__ void velocities(double rho, double mx, double my, double mz, double& u,
                                           double& v, double& w) {
    u = mx / rho;
    v = my / rho;
    w = mz / rho;
}

// HLLC flux through the face between L and R along DIR (0 x, 1 y, 2 z),
// hllc_riemann_solver_{x,y,z}_3d, hydro_solver_3d.c:239-416.  The branches of
// the reference become selects: the side (L or R) is chosen first, then the
// one physical flux and star state of that side are formed, so every lane of
// a wavefront runs the same instructions and each result is bit-identical to
// the branch the reference takes.
template <int DIR>
__device__ __forceinline__ void hllc(const Cell& L, const Cell& R, double gamma, double F[5]) {
    const double rL = rho_floor(L.d), rR = rho_floor(R.d);
    const double uL = DIR == 0 ? L.u : (DIR == 1 ? L.v : L.w);
    const double uR = DIR == 0 ? R.u : (DIR == 1 ? R.v : R.w);
    const double SL = tmin(uL - L.c, uR - R.c);
    const double SR = tmax(uL + L.c, uR + R.c);

    const double num = R.p - L.p + rL * uL * (SL - uL) - rR * uR * (SR - uR);
    const double den = rL * (SL - uL) - rR * (SR - uR);
    const double Ss = num / den;

    const bool sup_l = (SL >= 0.0);
    const bool sup_r = !sup_l && (SR <= 0.0);
    const bool left = sup_l || (!sup_r && (Ss >= 0.0));

    const double kd = left ? L.d : R.d;
    const double kmx = left ? L.mx : R.mx;
    const double kmy = left ? L.my : R.my;
    const double kmz = left ? L.mz : R.mz;
    const double ke = left ? L.e : R.e;
    const double ku = left ? L.u : R.u;
    const double kv = left ? L.v : R.v;
    const double kw = left ? L.w : R.w;
    const double kp = left ? L.p : R.p;
    const double kr = left ? rL : rR;
    const double kn = left ? uL : uR;  // normal velocity of the chosen side
    const double S = left ? SL : SR;

    // compute_flux_{x,y,z}_3d, hydro_solver_3d.c:199-233
    double f0, f1, f2, f3, f4;
    if (DIR == 0) {
        f0 = kmx; f1 = kmx * ku + kp; f2 = kmy * ku; f3 = kmz * ku;
    } else if (DIR == 1) {
        f0 = kmy; f1 = kmx * kv; f2 = kmy * kv + kp; f3 = kmz * kv;
    } else {
        f0 = kmz; f1 = kmx * kw; f2 = kmy * kw; f3 = kmz * kw + kp;
    }
    f4 = (ke + kp) * kn;

    // star state of the chosen side, hydro_solver_3d.c:267-273 / :283-289
    const double factor = (S - kn) / (S - Ss);
    const double sd = kr * factor;
    const double smx = sd * (DIR == 0 ? Ss : ku);
    const double smy = sd * (DIR == 1 ? Ss : kv);
    const double smz = sd * (DIR == 2 ? Ss : kw);
    const double se = sd * (ke / kr + (Ss - kn) * (Ss + kp / (kr * (S - kn))));

    const bool sup = sup_l || sup_r;
    F[0] = sup ? f0 : f0 + S * (sd - kd);
    F[1] = sup ? f1 : f1 + S * (smx - kmx);
    F[2] = sup ? f2 : f2 + S * (smy - kmy);
    F[3] = sup ? f3 : f3 + S * (smz - kmz);
    F[4] = sup ? f4 : f4 + S * (se - ke);
}

// The same flux with both cells resident in LDS as planes of doubles
// (field f of a cell at cell_address[f * FS]): the side is chosen from the
// normal velocity, pressure, sound speed and density of L and R, and only the
// chosen cell's fields are then read, through one selected address, instead
// of selecting every field between two register copies.  Bit-identical to
// hllc<DIR>.  Field order: d mx my mz e u v w p c.
template <int DIR, int FS>
__device__ __forceinline__ void hllc_pick(double dL, double unL, double pL, double cL, double dR,
                                          double unR, double pR, double cR,
                                          const double* __restrict__ aL,
                                          const double* __restrict__ aR, double gamma,
                                          double F[5]) {
    const double rL = rho_floor(dL), rR = rho_floor(dR);
    const double SL = tmin(unL - cL, unR - cR);
    const double SR = tmax(unL + cL, unR + cR);
    const double num = pR - pL + rL * unL * (SL - unL) - rR * unR * (SR - unR);
    const double den = rL * (SL - unL) - rR * (SR - unR);
    const double Ss = num / den;

    const bool sup_l = (SL >= 0.0);
    const bool sup_r = !sup_l && (SR <= 0.0);
    const bool left = sup_l || (!sup_r && (Ss >= 0.0));

    const double* k = left ? aL : aR;
    const double kd = k[0], kmx = k[FS], kmy = k[2 * FS], kmz = k[3 * FS], ke = k[4 * FS];
    const double ku = k[5 * FS], kv = k[6 * FS], kw = k[7 * FS], kp = k[8 * FS];
    const double kr = left ? rL : rR;
    const double kn = DIR == 0 ? ku : (DIR == 1 ? kv : kw);
    const double S = left ? SL : SR;

    double f0, f1, f2, f3, f4;
    if (DIR == 0) {
        f0 = kmx; f1 = kmx * ku + kp; f2 = kmy * ku; f3 = kmz * ku;
    } else if (DIR == 1) {
        f0 = kmy; f1 = kmx * kv; f2 = kmy * kv + kp; f3 = kmz * kv;
    } else {
        f0 = kmz; f1 = kmx * kw; f2 = kmy * kw; f3 = kmz * kw + kp;
    }
    f4 = (ke + kp) * kn;

    const double factor = (S - kn) / (S - Ss);
    const double sd = kr * factor;
    const double smx = sd * (DIR == 0 ? Ss : ku);
    const double smy = sd * (DIR == 1 ? Ss : kv);
    const double smz = sd * (DIR == 2 ? Ss : kw);
    const double se = sd * (ke / kr + (Ss - kn) * (Ss + kp / (kr * (S - kn))));

    const bool sup = sup_l || sup_r;
    F[0] = sup ? f0 : f0 + S * (sd - kd);
    F[1] = sup ? f1 : f1 + S * (smx - kmx);
    F[2] = sup ? f2 : f2 + S * (smy - kmy);
    F[3] = sup ? f3 : f3 + S * (smz - kmz);
    F[4] = sup ? f4 : f4 + S * (se - ke);
}

// ---------------------------------------------------------------------------
// double-double sums (compensated): two_sum is Knuth's error-free addition;
// dd_add is the accurate pair addition of Joldes, Muller and Popescu (2017,
// Algorithm 6).  dd_add is commutative bit for bit, so an xor butterfly over
// lanes reproduces a pairwise tree over consecutive leaves.
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {
    DD r;
    r.hi = a + b;
    double bb = r.hi - a;
    r.lo = (a - (r.hi - bb)) + (b - bb);
    return r;
}
__device__ __forceinline__ DD fast_two_sum(double a, double b) {
    DD r;
    r.hi = a + b;
    r.lo = b - (r.hi - a);
    return r;
}
__device__ __forceinline__ DD dd_add(DD x, DD y) {
    DD s = two_sum(x.hi, y.hi);
    DD t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    DD v = fast_two_sum(s.hi, c);
    double w = t.lo + v.lo;
    return fast_two_sum(v.hi, w);
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
    return __shfl_xor(v, m, 64);
}

// Whole-wavefront lane shifts by one lane as DPP moves (wave_shr:1 = 0x138,
// wave_shl:1 = 0x130; two 32-bit moves per double, no LDS round trip).  The
// lane with no source (lane 0 for shift_up, lane 63 for shift_down) keeps
// `edge`.  Every lane of the wavefront must be active.
__device__ __forceinline__ double lane_shift_up(double v, double edge) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(edge), __double2loint(v), 0x138,
                                               0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), __double2hiint(v), 0x138,
                                               0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_shift_down(double v, double edge) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(edge), __double2loint(v), 0x130,
                                               0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), __double2hiint(v), 0x130,
                                               0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum of a double-double over the 64 lanes of a wavefront as a pairwise tree
// over lane order (offsets 1, 2, 4, ... 32); every lane ends with the sum.
__device__ __forceinline__ DD wave_dd_tree(DD v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        DD p;
        p.hi = shfl_xor_d(v.hi, o);
        p.lo = shfl_xor_d(v.lo, o);
        v = dd_add(v, p);
    }
    return v;
}

// Max of a non-negative double over a wavefront, as ordered bit patterns.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long p = __shfl_xor(v, o, 64);
        v = p > v ? p : v;
    }
    return v;
}

// Max over a block (blockDim.x a multiple of 64, at most 1024 threads); the
// result is valid in thread 0.  Every thread of the block must call it.
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v) {
    __shared__ unsigned long long s_part[16];
    v = wave_max_u64(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int q = 1; q < nw; ++q) v = s_part[q] > v ? s_part[q] : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long dbits(double x) {
    return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

}  // namespace hyd
