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
// iteration t is accumulated in res[t & 3]; a launch clears the slots it is
// about to fill.  A launch of iteration t runs only while the residuals of
// iterations t-1 and t-2 are both >= tol; the first launch that sees one
// below it sets converged and iters = t, and the host works out which
// iteration met the tolerance (t-1, or t-2 when t-2 and t-1 ran as a pair).
struct SorCtl {
    unsigned long long res[4];  // max |phi_GS - phi_old| per iteration (ordered bits)
    int converged;
    int iters;                  // first iteration that did not run
    double last_res;
};

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
// Clear the residual slots of iterations it and it + 1 (unless converged).
hipError_t launch_sor_prepare(SorCtl* ctl, int it, hipStream_t st);
// The stop check of iteration it alone (after the last launch of a batch).
hipError_t launch_sor_check(SorCtl* ctl, int it, double tol, hipStream_t st);
// Two iterations in one pass, out of place: (in_r, in_b) -> (out_r, out_b)
// on the interior (boundary cells of out must already hold the boundary
// values).  With ctl, the residuals of iterations it and it + 1 go to their
// slots and the pass is skipped once an earlier iteration met tol.
hipError_t launch_sor2(const GridDims& g, const SorCoef& c, const double* in_r,
                       const double* in_b, double* out_r, double* out_b, const double* rhs_r,
                       const double* rhs_b, SorCtl* ctl, int it, double tol, hipStream_t st);
// The same for the FP32 correction equation L e = r.
hipError_t launch_sor2_f32(const GridDims& g, const SorCoef& c, const float* in_r,
                           const float* in_b, float* out_r, float* out_b, const float* rhs_r,
                           const float* rhs_b, hipStream_t st);

// g = -grad phi with the boundary copies of gravity_solver_3d.c:820-897.
hipError_t launch_acceleration(const GridDims& g, const double* phi_r, const double* phi_b,
                               double* gx, double* gy, double* gz, hipStream_t st);

// ---- defect correction ---------------------------------------------------
// r = f - L phi (FP64 arithmetic, FP32 store, r_r/r_b may be null) on
// interior cells; 0 on the boundary.  Writes sum r^2 and sum f^2 to acc[0],
// acc[1], summed per block into part (residual_scratch_len doubles) and then
// in a fixed order.
long residual_scratch_len(const GridDims& g);
hipError_t launch_residual(const GridDims& g, const SorCoef& c, const double* phi_r,
                           const double* phi_b, const double* rhs_r, const double* rhs_b,
                           float* r_r, float* r_b, double* part, double* acc, hipStream_t st);
// FP32 red-black sweep of L e = r, e = 0 on the boundary.
hipError_t launch_sor_sweep_f32(const GridDims& g, const SorCoef& c, int colour, float* e_r,
                                float* e_b, const float* r_r, const float* r_b, hipStream_t st);
// phi += e on interior cells (FP64 accumulation), e := 0.
hipError_t launch_correct(const GridDims& g, double* phi_r, double* phi_b, float* e_r,
                          float* e_b, hipStream_t st);

}  // namespace hyd
