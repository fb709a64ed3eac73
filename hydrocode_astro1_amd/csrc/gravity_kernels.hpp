// gravity_kernels.hpp -- launch interface of the Poisson / multipole kernels.
//
// Device layout of the potential: red-black split.  Cell (i,j,k) has colour
// c = (i+j+k)&1 and lives at arr[c][(i*ny + j)*nzh + (k>>1)], nzh = (nz+1)/2.
// A colour sweep then reads and writes whole cache lines of one array.
#pragma once

#include <hip/hip_runtime.h>

namespace hyd {

constexpr int MP_LMAX_MAX = 8;

struct GridDims {
    int nx, ny, nz, nzh;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double dx, dy, dz;
};

// Per-solve control of the SOR iteration on the device.  The residual of
// iteration t (the largest |phi_GS - phi_old|, as ordered bits) is the max
// over SOR_SLOTS slots of slot[t & 3]; a block adds its maximum to slot
// (block & (SOR_SLOTS - 1)), each slot on its own 128-byte line, so the
// atomics of one sweep spread over SOR_SLOTS addresses.  Iteration t runs
// only while the residual of iteration t-1 is >= tol: a check kernel before
// the two colour sweeps of t evaluates that rule, clears the slots of t+1,
// and on the first residual below tol sets converged and iters = t (the
// reference's iteration count, gravity_solver_3d.c:714-727).
constexpr int SOR_SLOTS = 256;
struct SorCtl {
    int converged;
    int iters;                  // first iteration that did not run
    double last_res;
    unsigned long long pad_[14];  // the header keeps a line of its own
    unsigned long long slot[4][SOR_SLOTS][16];  // value in [..][..][0]
};

// Host view: the residual of slot group q.
inline unsigned long long sor_slot_max(const SorCtl& c, int q) {
    unsigned long long m = 0ull;
    for (int s = 0; s < SOR_SLOTS; ++s) m = c.slot[q][s][0] > m ? c.slot[q][s][0] : m;
    return m;
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// Residual of slot group q, read by one whole wavefront (every lane gets it).
__device__ __forceinline__ unsigned long long sor_wave_slot_max(const SorCtl* c, int q) {
    const int lane = threadIdx.x & 63;
    unsigned long long v = 0ull;
    for (int s = lane; s < SOR_SLOTS; s += 64) {
        const unsigned long long w = c->slot[q][s][0];
        v = w > v ? w : v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// One block's maximum into its slot of group q (skipped when the slot
// already holds as much).
__device__ __forceinline__ void sor_slot_add(SorCtl* c, int q, unsigned block,
                                             unsigned long long v) {
    unsigned long long* p = &c->slot[q][block & (SOR_SLOTS - 1)][0];
    if (v > __atomic_load_n(p, __ATOMIC_RELAXED)) atomicMax(p, v);
}
#endif

struct SorCoef {
    double dx2i, dy2i, dz2i, factor, omega;
};

// ---- moments -------------------------------------------------------------
// Number of moment sums of pass 2: 9 Cartesian Q_ij, then A_lm, B_lm for
// l = 3..lmax, m = 0..l (layout [l][m][A,B]).
inline int mp_n_hi(int lmax) {
    int n = 0;
    for (int l = 3; l <= lmax; ++l) n += 2 * (l + 1);
    return n;
}
inline int mp_n_pass2(int lmax) { return 9 + mp_n_hi(lmax); }

// moments output layout (doubles): [0..3] M, xcm, ycm, zcm; [4..12] Q row
// major; [13..] hi-order A/B.  Scratch: partial double-doubles.
long moments_scratch_len(long ncells, int lmax);
hipError_t launch_moments(const GridDims& g, const double* rho, int lmax, double* scratch,
                          double* mom_out, hipStream_t st);
// The same sums in the reference's order (one cell after another): every
// term is stored (max(4, nq) x npad doubles), then summed by launch_seqsum
// with `method` (SeqSumMethod).  moments_seq_len gives the buffer length in
// doubles.
long moments_seq_npad(long ncells);
long moments_seq_len(long ncells, int lmax);
hipError_t launch_moments_sequential(const GridDims& g, const double* rho, int lmax, int method,
                                     double* buf, double* mom_out, hipStream_t st);

// ---- potential layout ----------------------------------------------------
hipError_t launch_natural_to_split(const GridDims& g, const double* nat, double scale,
                                   double* red, double* black, hipStream_t st);
hipError_t launch_split_to_natural(const GridDims& g, const double* red, const double* black,
                                   double* nat, hipStream_t st);

// ---- SOR -----------------------------------------------------------------
// One colour sweep of iteration `it` (colour 0 = red, 1 = black).  Exits at
// once when the previous iteration already met tol or the solve converged.
hipError_t launch_sor_sweep(const GridDims& g, const SorCoef& c, int colour, int it, double tol,
                            double* phi_r, double* phi_b, const double* rhs_r,
                            const double* rhs_b, SorCtl* ctl, hipStream_t st);
// Boundary values of the potential (isolated multipole or periodic copy),
// gravity_solver_3d.c:268-504.  bc: 0 periodic, 1 isolated, 2 dirichlet.
hipError_t launch_sor_boundary(const GridDims& g, int bc, double G, int lmax, const double* mom,
                               const double* wlm, double* phi_r, double* phi_b, hipStream_t st);
hipError_t launch_sor_reset(SorCtl* ctl, hipStream_t st);
// The stop check of iteration it alone (after the last launch of a batch).
hipError_t launch_sor_check(SorCtl* ctl, int it, double tol, hipStream_t st);
// g = -grad phi with the boundary copies of gravity_solver_3d.c:820-897.
hipError_t launch_acceleration(const GridDims& g, const double* phi_r, const double* phi_b,
                               double* gx, double* gy, double* gz, hipStream_t st);
// The 2D branch (nz = 1, gravity_solver_3d.c:569-634, :764-809): the
// potential in the natural layout, rhs = four_pi_G rho formed in the sweep.
hipError_t launch_sor2d_sweep(const GridDims& g, const SorCoef& c, int colour, int it, double tol,
                              double four_pi_G, double* phi, const double* rho, SorCtl* ctl,
                              hipStream_t st);
hipError_t launch_sor2d_boundary(const GridDims& g, int bc, double G, int lmax, const double* mom,
                                 const double* wlm, double* phi, hipStream_t st);
hipError_t launch_accel2d(const GridDims& g, const double* phi, double* gx, double* gy, double* gz,
                          hipStream_t st);

// ---- defect correction ---------------------------------------------------
// r = f - L phi (FP64 arithmetic, FP32 store, r_r/r_b may be null) on
// interior cells; 0 on the boundary.  Writes sum r^2 and sum f^2 to acc[0],
// acc[1], summed per block into part (residual_scratch_len doubles) and then
// in a fixed order.
long residual_scratch_len(const GridDims& g);
hipError_t launch_residual(const GridDims& g, const SorCoef& c, const double* phi_r,
                           const double* phi_b, const double* rhs_r, const double* rhs_b,
                           float* r_r, float* r_b, double* part, double* acc, hipStream_t st);
// acc[0] = sum over the interior of (r - L e)^2 (L e in FP32), acc[1] = 0;
// part as for launch_residual.
hipError_t launch_start_residual_sq(const GridDims& g, const SorCoef& c, const float* e_r,
                                   const float* e_b, const float* r_r, const float* r_b,
                                   double* part, double* acc, hipStream_t st);
// FP32 red-black sweep of L e = r, e = 0 on the boundary.
hipError_t launch_sor_sweep_f32(const GridDims& g, const SorCoef& c, int colour, float* e_r,
                                float* e_b, const float* r_r, const float* r_b, hipStream_t st);
// phi += e on interior cells (FP64 accumulation), e := 0.
hipError_t launch_correct(const GridDims& g, double* phi_r, double* phi_b, float* e_r,
                          float* e_b, hipStream_t st);

}  // namespace hyd
