// hydro_kernels.hip -- the Godunov step on the GPU.
//
// One step of godunov_step_3d (hydro_solver_3d.c:637-811) is two kernels:
//
//  hydro_step_kernel  reads the state planes U_old once, forms every HLLC face
//      flux on chip, applies the conservative update with floors to interior
//      cells and the gravity source to all cells, and writes U_pre: interior
//      cells into U_new, boundary cells into a small face buffer.  It also
//      reduces the signal speed of the updated interior for the next dt.
//  hydro_bc_kernel    writes the final boundary shell of U_new from U_pre in
//      closed form (periodic copies, then x, y, z boundary phases of
//      hydro_solver_3d.c:456-552 and :779-807) and reduces its signal speed.
//
// Layout: five SoA planes of doubles, flat index (i*ny + j)*nz + k, so k is
// the contiguous axis (IDX3D, hydro_solver_3d.h:12).
//
// hydro_step_kernel tiling.  A block owns an 8 (j) x 64 (k) column of cells
// and marches along x over a chunk of planes.  Ten wavefronts: waves 1..8 own
// rows j0..j0+7, one lane per k; wave 0 owns row j0-1 and wave 9 row j0+8,
// the y halo.  At each plane:
//   x faces: the cell of plane i+1 is carried in registers, F_x is formed by
//            the owning lane (no LDS), F_x(i-1) is kept from the last plane.
//   y faces: every row writes its cells (conserved + primitives) to LDS; row
//            r forms F_y(r, r+1) against row r+1 and passes it through LDS to
//            row r+1.
//   z faces: lane l forms F_z(l, l+1) against LDS; lane l+1 receives it by a
//            DPP lane shift.  The z halo cells k0-1 / k0+64 of rows 0..7 are
//            loaded by lanes 0..7 of the halo waves.
// Each face is formed once per block (plus the halo faces: 9 y rows for 8
// cell rows, 65 z faces for 64 cells), each cell's primitives once.
#include "device_common.hpp"
#include "hydro_kernels.hpp"

namespace hyd {

namespace {
// Rows per block: 10 cell rows + 2 halo rows = 12 wavefronts, 152 KiB of
// LDS, 3 waves per SIMD (measured fastest of 4, 6, 8, 10 at 256^3).
#ifndef HYD_TJ
#define HYD_TJ 10
#endif
constexpr int TK = 64;            // lanes (k) per row
constexpr int TJ = HYD_TJ;        // cell rows (j) per block
constexpr int NW = TJ + 2;        // wavefronts per block
constexpr int NT = NW * 64;       // threads per block
constexpr int LR = TJ + 2;        // LDS rows: j0-1 .. j0+TJ
constexpr int LC = TK + 2;        // LDS cols: k0-1 .. k0+64
constexpr int NF = 10;            // LDS fields per cell: d mx my mz e u v w p c
constexpr int FS = LR * LC;       // LDS stride between fields of one cell
}  // namespace

__device__ __forceinline__ double dt_from_speed(unsigned long long bits, unsigned long long nan,
                                                double dx, double dy, double dz, double cfl) {
    // compute_dt_3d, hydro_solver_3d.c:423 and :444-449.  After a NaN speed
    // the reference's running MAX restarts from the next speed, so its
    // 1e-10 start value no longer takes part (nan_fold_block).
    double ms = nan ? bitsd(bits) : tmax(1.0e-10, bitsd(bits));
    double tx = dx / ms, ty = dy / ms, tz = dz / ms;
    return cfl * tmin(tx, tmin(ty, tz));
}

struct Raw {
    double d, mx, my, mz, e;
};

__device__ __forceinline__ void load_raw(const HydroStepArgs& a, long idx, bool valid, Raw& r) {
    if (valid) {
        r.d = a.uin[0][idx];
        r.mx = a.uin[1][idx];
        r.my = a.uin[2][idx];
        r.mz = a.uin[3][idx];
        r.e = a.uin[4][idx];
    } else {  // outside the grid: a quiet placeholder whose fluxes are never used
        r.d = 1.0; r.mx = 0.0; r.my = 0.0; r.mz = 0.0; r.e = 1.0;
    }
}

__device__ __forceinline__ Cell to_cell(const Raw& r, double gamma) {
    Cell q;
    q.d = r.d; q.mx = r.mx; q.my = r.my; q.mz = r.mz; q.e = r.e;
    prims(q, gamma);
    return q;
}

typedef double LdsCells[NF][LR][LC];

__device__ __forceinline__ void lds_put(LdsCells& s, int r, int c, const Cell& q) {
    s[0][r][c] = q.d; s[1][r][c] = q.mx; s[2][r][c] = q.my; s[3][r][c] = q.mz;
    s[4][r][c] = q.e; s[5][r][c] = q.u; s[6][r][c] = q.v; s[7][r][c] = q.w;
    s[8][r][c] = q.p; s[9][r][c] = q.c;
}
__device__ __forceinline__ Cell lds_get(const LdsCells& s, int r, int c) {
    Cell q;
    q.d = s[0][r][c]; q.mx = s[1][r][c]; q.my = s[2][r][c]; q.mz = s[3][r][c];
    q.e = s[4][r][c]; q.u = s[5][r][c]; q.v = s[6][r][c]; q.w = s[7][r][c];
    q.p = s[8][r][c]; q.c = s[9][r][c];
    return q;
}

// HLLC flux between two cells that are both in an LDS plane buffer
template <int DIR>
__device__ __forceinline__ void hllc_lds(const LdsCells& s, int rl, int cl, int rr, int cr,
                                         double gamma, double F[5]) {
    constexpr int UN = 5 + DIR;  // field of the normal velocity
    const double* aL = &s[0][rl][cl];
    const double* aR = &s[0][rr][cr];
    hllc_pick<DIR, FS>(aL[0], aL[UN * FS], aL[8 * FS], aL[9 * FS], aR[0], aR[UN * FS], aR[8 * FS],
                       aR[9 * FS], aL, aR, gamma, F);
}

__device__ __forceinline__ long face_index(int nx, int ny, int nz, int i, int j, int k, int f) {
    // Boundary cell -> slot in the face buffer: x faces take cells with i on a
    // boundary, y faces the rest with j on a boundary, z faces the remainder.
    if (i == 0 || i == nx - 1) {
        int side = (i != 0);
        return ((long)(side * 5 + f) * ny + j) * nz + k;
    }
    long off = 10L * ny * nz;
    if (j == 0 || j == ny - 1) {
        int side = (j != 0);
        return off + ((long)(side * 5 + f) * nx + i) * nz + k;
    }
    off += 10L * nx * nz;
    int side = (k != 0);
    return off + ((long)(side * 5 + f) * nx + i) * ny + j;
}

// LDS coordinates: cell (row, col) with row in -1..TJ, col in -1..64 lives
// at [row+1][col+1] of the plane buffer; planes alternate between two
// buffers, so each plane needs two barriers: B1 (the plane is complete) and
// B2 (its F_y and the next plane are complete).  s_fy[r+1] holds F_y(r, r+1)
// for r = -1..TJ-1.  s_fzl[p][r] holds F_z(k0-1, k0) of row r for the plane
// of parity p; wave 0 forms it one plane ahead.
//
// The three roles (cell rows, the low halo row, the high halo row) run
// separate loops with the same barrier sequence, so each role's registers
// are allocated for that role alone.
typedef double PlaneBufs[2][NF][LR][LC];
typedef double FyRows[5][TJ + 1][TK];
typedef double FzHalo[2][5][TJ];

// One segment of a block's work: tile (kt, jt), planes ib .. ie-1.  smax and
// nan_speed accumulate the signal speed of the cell rows' updated interior
// over the segments of the block.
template <bool GRAV>
__device__ __forceinline__ void march_segment(const HydroStepArgs& a, PlaneBufs& s_cell,
                                              FyRows& s_fy, FzHalo& s_fzl, int kt, int jt,
                                              int ib, int ie, int wave, int lane, double dt,
                                              const EosConst& kc, unsigned long long& smax,
                                              bool& nan_speed) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int j0 = jt * TJ, k0 = kt * TK;

    const int row = wave - 1;  // -1 .. TJ
    const int j = j0 + row;
    const int k = k0 + lane;
    const bool valid = (j >= 0) && (j < ny) && (k < nz);

    // z halo cells: wave 0 lanes 0..TJ-1 load (j0+lane, k0-1), wave NW-1
    // lanes 0..TJ-1 load (j0+lane, k0+64).
    const bool has_halo = (wave == 0 || wave == NW - 1) && (lane < TJ);
    const int hj = j0 + lane;
    const int hk = (wave == 0) ? k0 - 1 : k0 + TK;
    const int hcol = (wave == 0) ? 0 : LC - 1;
    const bool hvalid = has_halo && (hj < ny) && (hk >= 0) && (hk < nz);

    const long sx = (long)ny * nz;
    const long cidx = ((long)ib * ny + j) * nz + k;  // own cell at plane ib
    const long hidx = ((long)ib * ny + hj) * nz + hk;
    const double gamma = kc.gamma;

    // prologue: plane ib into buffer ib & 1
    {
        Raw r;
        load_raw(a, cidx, valid, r);
        lds_put(s_cell[ib & 1], row + 1, lane + 1, to_cell(r, gamma));
        if (has_halo) {
            Raw h;
            load_raw(a, hidx, hvalid, h);
            lds_put(s_cell[ib & 1], lane + 1, hcol, to_cell(h, gamma));
        }
    }
    __syncthreads();

    if (wave >= 1 && wave <= TJ) {
        // ---------------------------------------------------- cell rows
        const double ax = dt / a.dx, ay = dt / a.dy, az = dt / a.dz;
        // block-uniform: this tile holds the z-min face (lane 0) or the
        // z-max face (lane nz-1-k0, with its neighbour in the same wave)
        const bool zlo = a.zcopy && k0 == 0;
        const bool zhi = a.zcopy && k0 + TK >= nz;
        double fxl[5] = {0, 0, 0, 0, 0};
        if (ib > 0) {  // F_x(ib-1) for the update of plane ib
            Raw pr;
            load_raw(a, cidx - sx, valid, pr);
            hllc<0>(to_cell(pr, gamma), lds_get(s_cell[ib & 1], row + 1, lane + 1), gamma, fxl);
        }
        // plane i+1 is loaded one plane ahead of its use, so its latency has a
        // whole plane of work to hide behind
        Raw rn;
        if (ib + 1 < nx) load_raw(a, cidx + sx, valid, rn);
        for (int i = ib; i < ie; ++i) {
            const long off = (long)(i - ib) * sx;
            const bool has_next = (i + 1 < nx);
            const int pb = i & 1;
            __syncthreads();  // B1: plane i complete in LDS
            Raw rn2;  // plane i+2, used at the next plane
            if (i + 2 < nx && i + 1 < ie) load_raw(a, cidx + off + 2 * sx, valid, rn2);

            const Cell cur = lds_get(s_cell[pb], row + 1, lane + 1);
            const double* me = &s_cell[pb][0][row + 1][lane + 1];
            {
                double fy[5];
                const double* nb = &s_cell[pb][0][row + 2][lane + 1];
                hllc_pick<1, FS>(cur.d, cur.v, cur.p, cur.c, nb[0], nb[6 * FS], nb[8 * FS],
                                 nb[9 * FS], me, nb, gamma, fy);
#pragma unroll
                for (int f = 0; f < 5; ++f) s_fy[f][row + 1][lane] = fy[f];
            }
            double dz[5];
            {
                double fz[5];
                const double* nb = &s_cell[pb][0][row + 1][lane + 2];
                hllc_pick<2, FS>(cur.d, cur.w, cur.p, cur.c, nb[0], nb[7 * FS], nb[8 * FS],
                                 nb[9 * FS], me, nb, gamma, fz);
                // F_z(k-1, k) from the lane below; lane 0 takes the halo
                // face formed by the low halo wave
#pragma unroll
                for (int f = 0; f < 5; ++f) dz[f] = fz[f] - lane_shift_up(fz[f], s_fzl[pb][f][row]);
            }
            double tx[5] = {0, 0, 0, 0, 0};
            if (has_next) {
                const Cell nxt = to_cell(rn, gamma);
                lds_put(s_cell[pb ^ 1], row + 1, lane + 1, nxt);
                double fxr[5];
                hllc_pick<0, FS>(cur.d, cur.u, cur.p, cur.c, nxt.d, nxt.u, nxt.p, nxt.c, me,
                                 &s_cell[pb ^ 1][0][row + 1][lane + 1], gamma, fxr);
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    tx[f] = -ax * (fxr[f] - fxl[f]);
                    fxl[f] = fxr[f];
                }
            }
            __syncthreads();  // B2: F_y of plane i and plane i+1 complete

            if (valid) {
                double q[5] = {cur.d, cur.mx, cur.my, cur.mz, cur.e};
                const bool interior = (i >= 1) && (i <= nx - 2) && (j >= 1) && (j <= ny - 2) &&
                                      (k >= 1) && (k <= nz - 2);
                if (interior) {
                    // hydro_solver_3d.c:738-762: x, then y, then z, left to right
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        double t = tx[f] - ay * (s_fy[f][row + 1][lane] - s_fy[f][row][lane]);
                        t = t - az * dz[f];
                        q[f] = q[f] + t;
                    }
                    floor_roundtrip(q[0], q[1], q[2], q[3], q[4], kc);
                }
                const long idx = cidx + off;
                if (GRAV) {
                    // apply_gravity_source_terms, hydro_solver_3d.c:606-627
                    const double gx = a.g[0][idx], gy = a.g[1][idx], gz = a.g[2][idx];
                    const double rho = q[0];
                    double u, v, w;
                    velocities(rho, q[1], q[2], q[3], u, v, w);
                    q[1] = q[1] + dt * rho * gx;
                    q[2] = q[2] + dt * rho * gy;
                    q[3] = q[3] + dt * rho * gz;
                    const double work = rho * (u * gx + v * gy + w * gz);
                    q[4] = q[4] + dt * work;
                }
                // z faces with outflow on both sides: the final value of
                // (i, j, 0) is that of (i, j, 1) and of (i, j, nz-1) that of
                // (i, j, nz-2), the neighbouring lane's new state
                // (hydro_solver_3d.c:535-550 after the update and gravity
                // source).  Its signal speed equals the neighbour's, which is
                // already in the maximum.
                double qz[5];
                if (zlo) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) qz[f] = lane_shift_down(q[f], q[f]);
                }
                if (zhi) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        const double t = lane_shift_up(q[f], q[f]);
                        qz[f] = (k == nz - 1) ? t : qz[f];
                    }
                }
                if (interior) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) a.uout[f][idx] = q[f];
                    // the interior density is at or above the floor here
                    unsigned long long sb =
                        dbits(signal_speed_floored(q[0], q[1], q[2], q[3], q[4], kc, nan_speed));
                    smax = sb > smax ? sb : smax;
                } else if (a.zcopy && (k == 0 || k == nz - 1) && i >= 1 && i <= nx - 2 &&
                           j >= 1 && j <= ny - 2) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) a.uout[f][idx] = qz[f];
                } else {
#pragma unroll
                    for (int f = 0; f < 5; ++f) a.face[face_index(nx, ny, nz, i, j, k, f)] = q[f];
                }
            }
            rn = rn2;
        }
    } else if (wave == 0) {
        // ------------------------------------------- low halo row (j0-1)
        // F_y(j0-1, j0) each plane; loads the left z halo and forms
        // F_z(k0-1, k0) of every cell row one plane ahead.
        if (lane < TJ) {
            double fh[5];
            hllc_lds<2>(s_cell[ib & 1], lane + 1, 0, lane + 1, 1, gamma, fh);
#pragma unroll
            for (int f = 0; f < 5; ++f) s_fzl[ib & 1][f][lane] = fh[f];
        }
        Raw rn, rh;  // plane i+1, loaded one plane ahead
        if (ib + 1 < nx) {
            load_raw(a, cidx + sx, valid, rn);
            if (has_halo) load_raw(a, hidx + sx, hvalid, rh);
        }
        for (int i = ib; i < ie; ++i) {
            const long off = (long)(i - ib) * sx;
            const bool has_next = (i + 1 < nx);
            const int pb = i & 1;
            __syncthreads();  // B1
            Raw rn2, rh2;
            if (i + 2 < nx && i + 1 < ie) {
                load_raw(a, cidx + off + 2 * sx, valid, rn2);
                if (has_halo) load_raw(a, hidx + off + 2 * sx, hvalid, rh2);
            }
            {
                double fy[5];
                hllc_lds<1>(s_cell[pb], 0, lane + 1, 1, lane + 1, gamma, fy);
#pragma unroll
                for (int f = 0; f < 5; ++f) s_fy[f][0][lane] = fy[f];
            }
            if (has_next) {
                lds_put(s_cell[pb ^ 1], 0, lane + 1, to_cell(rn, gamma));
                if (has_halo) lds_put(s_cell[pb ^ 1], lane + 1, 0, to_cell(rh, gamma));
            }
            __syncthreads();  // B2: plane i+1 complete
            if (has_next && lane < TJ) {
                double fh[5];
                hllc_lds<2>(s_cell[pb ^ 1], lane + 1, 0, lane + 1, 1, gamma, fh);
#pragma unroll
                for (int f = 0; f < 5; ++f) s_fzl[pb ^ 1][f][lane] = fh[f];
            }
            rn = rn2;
            rh = rh2;
        }
    } else {
        // ------------------------------------------ high halo row (j0+TJ)
        // loads row j0+TJ and the right z halo of the next plane.
        Raw rn, rh;  // plane i+1, loaded one plane ahead
        if (ib + 1 < nx) {
            load_raw(a, cidx + sx, valid, rn);
            if (has_halo) load_raw(a, hidx + sx, hvalid, rh);
        }
        for (int i = ib; i < ie; ++i) {
            const long off = (long)(i - ib) * sx;
            const bool has_next = (i + 1 < nx);
            const int pb = i & 1;
            __syncthreads();  // B1
            Raw rn2, rh2;
            if (i + 2 < nx && i + 1 < ie) {
                load_raw(a, cidx + off + 2 * sx, valid, rn2);
                if (has_halo) load_raw(a, hidx + off + 2 * sx, hvalid, rh2);
            }
            if (has_next) {
                lds_put(s_cell[pb ^ 1], LR - 1, lane + 1, to_cell(rn, gamma));
                if (has_halo) lds_put(s_cell[pb ^ 1], lane + 1, LC - 1, to_cell(rh, gamma));
            }
            __syncthreads();  // B2
            rn = rn2;
            rh = rh2;
        }
    }
}

// Block -> work.  The (tile, plane) pairs are taken in one flat order, tile
// after tile (k tiles fastest, then j tiles) and x fastest within a tile;
// each block takes the next a.wpb of them, which may end one tile and start
// the next (two segments, each with its own prologue).  With as many blocks
// as CUs every CU gets the same number of planes in a single round.
// Consecutive blocks land on different XCDs; they are regrouped so that each
// XCD walks a contiguous run of tiles (shared halos stay in its L2).
template <bool GRAV>
__global__ void __launch_bounds__(NT) hydro_step_kernel(HydroStepArgs a) {
    __shared__ PlaneBufs s_cell;
    __shared__ FyRows s_fy;
    __shared__ FzHalo s_fzl;
    __shared__ unsigned long long s_smax[NW];

    const int nx = a.nx;
    // wavefront index made scalar so role branches are uniform jumps
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;

    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    const long total = (long)a.ntk * a.ntj * nx;
    long f = (long)b * a.wpb;
    const long fend = min(total, f + (long)a.wpb);

    const double dt = dt_from_speed(*a.ms_in, *a.nan_in, a.dx, a.dy, a.dz, a.cfl);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.dt_out = dt;
        *a.ms_clear = 0ull;
        *a.nan_clear = 0ull;
    }
    const EosConst kc = eos_const(a.gamma);
    unsigned long long smax = 0ull;
    bool nan_speed = false;  // a NaN signal speed among this lane's new cells
    while (f < fend) {
        const int t = (int)(f / nx);
        const int ib = (int)(f - (long)t * nx);
        const int ie = (int)min((long)nx, ib + (fend - f));
        march_segment<GRAV>(a, s_cell, s_fy, s_fzl, t % a.ntk, t / a.ntk, ib, ie, wave, lane, dt, kc,
                      smax, nan_speed);
        f += ie - ib;
        // the low halo wave may still read the last plane buffer
        __syncthreads();
    }
    if (wave >= 1 && wave <= TJ) s_smax[wave] = wave_max_u64(smax);
    // one atomic per block for the signal speed of the updated interior; a
    // NaN speed only raises the flag (hydro_nan_fold_kernel, after the
    // boundary kernel, then reduces the whole state in the reference's order)
    if (nan_speed) *a.nan_out = 1ull;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0ull;
        for (int w = 1; w <= TJ; ++w) m = s_smax[w] > m ? s_smax[w] : m;
        if (m != 0ull && m > __atomic_load_n(a.ms_out, __ATOMIC_RELAXED)) atomicMax(a.ms_out, m);
    }
}

// compute_dt_3d's reduction (hydro_solver_3d.c:423-441) when a speed is NaN,
// by one block over the whole state: the reference folds the speeds of cell
// after cell (u, v, w each) with MAX(a, b) = a > b ? a : b, so a NaN speed
// replaces the running value and the next speed replaces a NaN.  The result
// is NaN when the last speed is NaN, else the largest speed after the last
// NaN one (the 1e-10 start value drops out).  Writes it to *ms_out and sets
// the flag to 2; leaves both alone when the state has no NaN speed.
__device__ void nan_fold_block(const double* const u[5], long n, double gamma,
                               unsigned long long* ms_out, unsigned long long* nan_flag) {
    __shared__ unsigned long long s_last;
    unsigned long long last = 0ull;  // position + 1 of the last NaN speed, position 3 c + q
    for (long c = threadIdx.x; c < n; c += blockDim.x) {
        double s[3];
        cell_speeds(u[0][c], u[1][c], u[2][c], u[3][c], u[4][c], gamma, s);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (s[q] != s[q]) last = 3ull * (unsigned long long)c + q + 1ull;
    }
    last = block_max_u64(last);
    if (threadIdx.x == 0) s_last = last;
    __syncthreads();
    last = s_last;
    if (last == 0ull) return;  // uniform across the block
    unsigned long long m = 0ull;
    for (long c = (long)(last / 3) + threadIdx.x; c < n; c += blockDim.x) {
        double s[3];
        cell_speeds(u[0][c], u[1][c], u[2][c], u[3][c], u[4][c], gamma, s);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const unsigned long long pos = 3ull * (unsigned long long)c + q;
            if (pos >= last) {
                const unsigned long long b = dbits(s[q]);
                m = b > m ? b : m;
            }
        }
    }
    __syncthreads();  // s_part of block_max_u64 is reused
    m = block_max_u64(m);
    if (threadIdx.x == 0) {
        *ms_out = (last == 3ull * (unsigned long long)n) ? dbits(__builtin_nan("")) : m;
        *nan_flag = 2ull;
    }
}

// ---------------------------------------------------------------------------
// Boundary shell.  For a boundary cell c the reference's sequence
//   periodic x, periodic y, periodic z, then x, y, z boundary phases
// resolves to: take U_pre at the source cell s (per axis: an outflow side
// moves the coordinate one cell inward, a periodic pair maps 0 to n-1), then
// apply the reflections of the axes on which c sits at a reflective side,
// x first.  Sources are read from the face buffer when they are boundary
// cells and from U_new (final interior values) otherwise, so the in-place
// writes below never feed another thread's read.

template <int DIR>
__device__ __forceinline__ void maybe_reflect(bool on, double q[5], double gamma) {
    if (on) reflect<DIR>(q[0], q[1], q[2], q[3], q[4], gamma);
}

// One thread per boundary cell: blockIdx.z is the face (x-min, x-max,
// y-min, y-max, z-min, z-max), blockIdx.y the outer index and the threads of
// blockIdx.x the inner one (k on the x and y faces, j on the z faces).  Each
// face owns the cells not owned by an earlier axis: x faces all of theirs, y
// faces i in 1..nx-2, z faces i, j in 1..nx-2, 1..ny-2.
__global__ void __launch_bounds__(256) hydro_bc_kernel(HydroBCArgs a) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int face = blockIdx.z, axis = face >> 1, hi_side = face & 1;
    const int outer = blockIdx.y, inner = blockIdx.x * 256 + threadIdx.x;
    unsigned long long smax = 0ull;
    int i = 0, j = 0, k = 0;
    bool on = false;
    if (axis == 0) {
        i = hi_side ? nx - 1 : 0; j = outer; k = inner;
        on = (j < ny) && (k < nz);
    } else if (axis == 1) {
        i = 1 + outer; j = hi_side ? ny - 1 : 0; k = inner;
        on = (i <= nx - 2) && (k < nz);
    } else {
        i = 1 + outer; j = 1 + inner; k = hi_side ? nz - 1 : 0;
        on = (i <= nx - 2) && (j <= ny - 2);
    }
    if (on) {
        const int n3[3] = {nx, ny, nz};
        const int c3[3] = {i, j, k};
        int s3[3];
        bool refl[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int lo = a.bc[2 * d], hi = a.bc[2 * d + 1];
            const bool both_per = (lo == 0) && (hi == 0);  // BC_PERIODIC == 0
            int s = c3[d];
            bool r = false;
            if (c3[d] == 0) {
                if (lo == 2) s = 1;                // outflow
                else if (both_per) s = n3[d] - 1;  // periodic copy
                r = (lo == 1);
            } else if (c3[d] == n3[d] - 1) {
                if (hi == 2) s = n3[d] - 2;
                r = (hi == 1);
            }
            s3[d] = s;
            refl[d] = r;
        }
        double q[5];
        const bool src_boundary = (s3[0] == 0 || s3[0] == nx - 1 || s3[1] == 0 ||
                                   s3[1] == ny - 1 || s3[2] == 0 || s3[2] == nz - 1);
        if (src_boundary) {
#pragma unroll
            for (int f = 0; f < 5; ++f) q[f] = a.face[face_index(nx, ny, nz, s3[0], s3[1], s3[2], f)];
        } else {
            const long sidx = ((long)s3[0] * ny + s3[1]) * nz + s3[2];
#pragma unroll
            for (int f = 0; f < 5; ++f) q[f] = a.u[f][sidx];
        }
        maybe_reflect<0>(refl[0], q, a.gamma);
        maybe_reflect<1>(refl[1], q, a.gamma);
        maybe_reflect<2>(refl[2], q, a.gamma);
        const long idx = ((long)i * ny + j) * nz + k;
#pragma unroll
        for (int f = 0; f < 5; ++f) a.u[f][idx] = q[f];
        bool nan = false;
        const unsigned long long sb =
            dbits(signal_speed(q[0], q[1], q[2], q[3], q[4], a.gamma, nan));
        smax = sb > smax ? sb : smax;
        if (nan) *a.nan_out = 1ull;
    }
    smax = block_max_u64(smax);
    if (threadIdx.x == 0 && smax != 0ull && smax > __atomic_load_n(a.ms_out, __ATOMIC_RELAXED))
        atomicMax(a.ms_out, smax);
}

// Signal speed of a whole state (after an upload).
__global__ void __launch_bounds__(256) hydro_maxspeed_kernel(const double* __restrict__ d,
                                                             const double* __restrict__ mx,
                                                             const double* __restrict__ my,
                                                             const double* __restrict__ mz,
                                                             const double* __restrict__ e,
                                                             long n, double gamma,
                                                             unsigned long long* out,
                                                             unsigned long long* nan_flag) {
    unsigned long long smax = 0ull;
    bool nan = false;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long)gridDim.x * blockDim.x) {
        unsigned long long sb = dbits(signal_speed(d[c], mx[c], my[c], mz[c], e[c], gamma, nan));
        smax = sb > smax ? sb : smax;
    }
    if (nan) *nan_flag = 1ull;
    smax = block_max_u64(smax);
    if (threadIdx.x == 0 && smax != 0ull && smax > __atomic_load_n(out, __ATOMIC_RELAXED))
        atomicMax(out, smax);
}

// One block: the reference-order reduction when a NaN speed was flagged
// (by hydro_maxspeed_kernel, or by the step and boundary kernels of a step,
// launched after them so that it sees every cell's final value).
__global__ void __launch_bounds__(256) hydro_nan_fold_kernel(const double* d, const double* mx,
                                                             const double* my, const double* mz,
                                                             const double* e, long n, double gamma,
                                                             unsigned long long* out,
                                                             unsigned long long* nan_flag) {
    if (*nan_flag == 0ull) return;
    const double* u[5] = {d, mx, my, mz, e};
    nan_fold_block(u, n, gamma, out, nan_flag);
}

// AoS (host layout, 5 doubles per cell) <-> SoA planes.
__global__ void aos_to_soa_kernel(const double* __restrict__ aos, double* d, double* mx,
                                  double* my, double* mz, double* e, long n) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long)gridDim.x * blockDim.x) {
        d[c] = aos[5 * c]; mx[c] = aos[5 * c + 1]; my[c] = aos[5 * c + 2];
        mz[c] = aos[5 * c + 3]; e[c] = aos[5 * c + 4];
    }
}
__global__ void soa_to_aos_kernel(double* __restrict__ aos, const double* d, const double* mx,
                                  const double* my, const double* mz, const double* e, long n) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long)gridDim.x * blockDim.x) {
        aos[5 * c] = d[c]; aos[5 * c + 1] = mx[c]; aos[5 * c + 2] = my[c];
        aos[5 * c + 3] = mz[c]; aos[5 * c + 4] = e[c];
    }
}

// ---------------------------------------------------------------- launchers
// (tile, plane) pairs per block when the caller leaves it open.  A block
// fills a CU's LDS, so the grid runs in rounds of one block per CU; a block
// pays about one plane more per segment it starts (the plane before its
// first one, for F_x).  Choose the number of rounds R (blocks = R x CUs, the
// pairs split evenly) that minimises R x (pairs per block + 2); at least 8
// pairs per block.
static int choose_wpb(long total) {
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            v > 0)
            ncu = v;
        else
            ncu = 256;
    }
    long best_w = total, best = -1;
    for (int r = 1; r <= 8; ++r) {
        const long w = (total + (long)r * ncu - 1) / ((long)r * ncu);
        if (w < 8 && r > 1) break;
        const long cost = (long)r * (w + 2);
        if (best < 0 || cost < best) {
            best = cost;
            best_w = w;
        }
    }
    return (int)(best_w < 8 ? (total < 8 ? total : 8) : best_w);
}

hipError_t launch_hydro_step(const HydroStepArgs& a, hipStream_t st) {
    const int ntj = (a.ny + TJ - 1) / TJ, ntk = (a.nz + TK - 1) / TK;
    HydroStepArgs b = a;
    const long total = (long)ntj * ntk * a.nx;
    b.ntj = ntj; b.ntk = ntk;
    b.wpb = a.cx > 0 ? a.cx : choose_wpb(total);
    const long nblk = (total + b.wpb - 1) / b.wpb;
    // the instantiation without gravity carries no source-term code: 166
    // VGPRs and no spill, against 168 and three spilled for the other
    if (b.gravity)
        hipLaunchKernelGGL(hydro_step_kernel<true>, dim3((unsigned)nblk), dim3(NT), 0, st, b);
    else
        hipLaunchKernelGGL(hydro_step_kernel<false>, dim3((unsigned)nblk), dim3(NT), 0, st, b);
    return hipGetLastError();
}

bool hydro_zcopy(int nz, int bc_zmin, int bc_zmax) {
    return bc_zmin == 2 && bc_zmax == 2 && nz >= 3 && (nz - 1) % TK != 0;  // BC_OUTFLOW == 2
}

hipError_t launch_hydro_bc(const HydroBCArgs& a, hipStream_t st) {
    // widest inner and outer extents over the face kinds launched
    const int inner = a.zcopy ? a.nz : (a.nz > a.ny - 2 ? a.nz : a.ny - 2);
    const int outer = a.ny > a.nx - 2 ? a.ny : a.nx - 2;
    hipLaunchKernelGGL(hydro_bc_kernel,
                       dim3((unsigned)((inner + 255) / 256), (unsigned)outer, a.zcopy ? 4 : 6),
                       dim3(256), 0, st, a);
    // returns at once unless a NaN speed was flagged (a block-count ticket
    // in the boundary kernel for this cost 0.1 ms per step: 1024 blocks
    // queueing on one atomic)
    hipLaunchKernelGGL(hydro_nan_fold_kernel, dim3(1), dim3(256), 0, st, a.u[0], a.u[1], a.u[2],
                       a.u[3], a.u[4], (long)a.nx * a.ny * a.nz, a.gamma, a.ms_out, a.nan_out);
    return hipGetLastError();
}

hipError_t launch_hydro_maxspeed(double* const u[5], long n, double gamma, unsigned long long* out,
                                 unsigned long long* nan_flag, hipStream_t st) {
    long nb = (n + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(hydro_maxspeed_kernel, dim3((unsigned)nb), dim3(256), 0, st, u[0], u[1],
                       u[2], u[3], u[4], n, gamma, out, nan_flag);
    hipLaunchKernelGGL(hydro_nan_fold_kernel, dim3(1), dim3(256), 0, st, u[0], u[1], u[2], u[3],
                       u[4], n, gamma, out, nan_flag);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const double* aos, double* const u[5], long n, hipStream_t st) {
    long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3((unsigned)nb), dim3(256), 0, st, aos, u[0], u[1],
                       u[2], u[3], u[4], n);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(double* aos, double* const u[5], long n, hipStream_t st) {
    long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)nb), dim3(256), 0, st, aos, u[0], u[1],
                       u[2], u[3], u[4], n);
    return hipGetLastError();
}

}  // namespace hyd
