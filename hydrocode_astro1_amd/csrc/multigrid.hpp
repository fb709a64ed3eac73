// multigrid.hpp -- FP32 multigrid V-cycle for the correction equation of
// the defect-correction Poisson solve: L_h e = r with e = 0 on the boundary.
//
// Level 0 is the solver's own grid in the red-black split layout (fine cell
// x at x h, boundary cells 0 and n-1).  Coarse level l >= 1 has points at
// X_l(0) = 0 and X_l(c) = c 2^l - 1 (c >= 1) in fine cells, natural layout (i
// slowest), FP32: its end points are the fine boundary cells, its spacing
// H = 2^l except the first interval, H - 1.  Level l+1 keeps the even points
// of level l (that needs an even fine count n, then even interval counts).
// Operators are the vertex-centred 7-point Laplacian of these points (the
// non-uniform three-point form at index 1), smoothing red-black
// Gauss-Seidel/SOR, prolongation linear in position along each direction and
// restriction its transpose weighted by the dual-cell lengths, so a
// constant restricts to itself.
#pragma once
#include <hip/hip_runtime.h>

#include "gravity_kernels.hpp"

namespace hyd {

// One direction of a coarse level.
struct MgAxis {
    float a;         // 1 / H^2 (H in length units) at indices >= 2
    float am1, ap1;  // index 1: weights of its boundary-side and far neighbours
    float pw1;       // prolongation into index 1 from the next coarser level's point 1
    float r1[3];     // restriction into point 1 from the three finer points around it
};

struct MgLevel {
    int nx, ny, nz;  // points, both end points included
    MgAxis ax[3];    // x, y, z
};

// The coarse levels from one level down, run by a single workgroup in LDS
// (the levels' u, f, q at the given float offsets of its LDS buffer).
constexpr int MG_TAIL_MAX = 8;
constexpr size_t MG_TAIL_LDS_MAX = 150 * 1024;
struct MgTail {
    int count;  // levels in the table; the last is the coarsest
    int nu1, nu2;
    float omega;
    MgLevel lv[MG_TAIL_MAX];
    int off_u[MG_TAIL_MAX], off_f[MG_TAIL_MAX], off_q[MG_TAIL_MAX];
};

// level 0 (split) <-> level 1
hipError_t launch_mg_fine_residual(const GridDims& g, const SorCoef& c, const float* e_r,
                                   const float* e_b, const float* r_r, const float* r_b,
                                   float* q_r, float* q_b, hipStream_t st);
hipError_t launch_mg_restrict_fine(const GridDims& g, const float* q_r, const float* q_b,
                                   const MgLevel& c1, float* f1, float* u1, hipStream_t st);
hipError_t launch_mg_prolong_fine(const GridDims& g, const MgLevel& c1, const float* u1,
                                  float* e_r, float* e_b, hipStream_t st);

// coarse levels
hipError_t launch_mg_smooth(const MgLevel& L, int colour, float omega, float* u, const float* f,
                            hipStream_t st);
hipError_t launch_mg_residual(const MgLevel& L, const float* u, const float* f, float* q,
                              hipStream_t st);
hipError_t launch_mg_restrict(const MgLevel& fine, const float* q, const MgLevel& coarse,
                              float* fc, float* uc, hipStream_t st);
hipError_t launch_mg_prolong(const MgLevel& coarse, const float* uc, const MgLevel& fine,
                             float* uf, hipStream_t st);
// d_tab: an MgTail in device memory; f_first / u_first: the first level's
// f (read) and u (written) in global memory.
hipError_t launch_mg_tail(const MgTail* d_tab, int lds_floats, const float* f_first,
                          float* u_first, hipStream_t st);

}  // namespace hyd
