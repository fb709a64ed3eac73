// gravity_kernels.hip -- red-black SOR, multipole boundary potential,
// acceleration and the FP32 defect-correction pieces.
//
// The potential is stored red-black split (see gravity_kernels.hpp): a colour
// sweep reads the other colour's array and writes whole lines of its own.
// Red-black SOR is order independent within a colour, so each sweep matches
// the reference's sequential sweep (gravity_solver_3d.c:642-708) bit for bit.
#include "device_common.hpp"
#include "gravity_kernels.hpp"
#include "seqsum.hpp"

namespace hyd {

__device__ __forceinline__ long sidx(const GridDims& g, int i, int j, int k) {
    return ((long)i * g.ny + j) * g.nzh + (k >> 1);
}
__device__ __forceinline__ double phi_at(const GridDims& g, const double* r, const double* b,
                                         int i, int j, int k) {
    return (((i + j + k) & 1) ? b : r)[sidx(g, i, j, k)];
}

// ------------------------------------------------------------- layouts --
__global__ void natural_to_split_kernel(GridDims g, const double* __restrict__ nat, double scale,
                                        double* __restrict__ red, double* __restrict__ black) {
    const long n = (long)g.nx * g.ny * g.nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long ij = t / g.nz;
        const int j = (int)(ij % g.ny), i = (int)(ij / g.ny);
        double* dst = ((i + j + k) & 1) ? black : red;
        dst[sidx(g, i, j, k)] = scale * nat[t];
    }
}

__global__ void split_to_natural_kernel(GridDims g, const double* __restrict__ red,
                                        const double* __restrict__ black,
                                        double* __restrict__ nat) {
    const long n = (long)g.nx * g.ny * g.nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long ij = t / g.nz;
        const int j = (int)(ij % g.ny), i = (int)(ij / g.ny);
        nat[t] = phi_at(g, red, black, i, j, k);
    }
}

// ----------------------------------------------------------------- SOR --
// gravity_solver_3d.c:642-708 for one colour.  A block is 64 split entries
// (m) of 4 interior rows j of one plane i (blockIdx.z): lane m's cell is
// k = 2m + ((i+j+colour)&1).  No integer division in the index math.
constexpr int SOR_ROWS = 4;

// True when iteration `it` must not run; `writer` records the stop.  Called
// by every lane of one wavefront.
__device__ __forceinline__ bool sor_stop_check(SorCtl* ctl, int it, double tol, bool writer) {
    if (ctl->converged) return true;
    double prev = 0.0;
    bool stop = false;
    if (it > 0) {
        prev = bitsd(sor_wave_slot_max(ctl, (it - 1) & 3));
        stop = prev < tol;
    }
    if (stop && writer) {
        ctl->converged = 1;
        ctl->iters = it;
        ctl->last_res = prev;
    }
    return stop;
}

__global__ void __launch_bounds__(256) sor_sweep_kernel(GridDims g, SorCoef c, int colour, int it,
                                                        double tol, double* __restrict__ mine,
                                                        const double* __restrict__ other,
                                                        const double* __restrict__ rhs,
                                                        SorCtl* ctl) {
    // stop rule of gravity_solver_3d.c:715: sor_check_kernel, launched before
    // the red sweep of each iteration, sets converged when the iteration
    // must not run (tol < 0: no stop rule and no residual, the
    // relative-residual mode)
    if (tol >= 0.0 && ctl->converged) return;
    const unsigned blin = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;

    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * SOR_ROWS + (threadIdx.x >> 6);
    const int i = 1 + blockIdx.z;
    const int k = 2 * m + ((i + j + colour) & 1);
    unsigned long long rb = 0ull;
    if (j <= g.ny - 2 && k >= 1 && k <= g.nz - 2) {
        const long base = ((long)i * g.ny + j) * g.nzh;
        const long sxh = (long)g.ny * g.nzh;
        const double o_im = other[base - sxh + m], o_ip = other[base + sxh + m];
        const double o_jm = other[base - g.nzh + m], o_jp = other[base + g.nzh + m];
        const double o_km = other[base + ((k - 1) >> 1)], o_kp = other[base + ((k + 1) >> 1)];
        const double old = mine[base + m];
        const double lap = (o_im + o_ip) * c.dx2i + (o_jm + o_jp) * c.dy2i + (o_km + o_kp) * c.dz2i;
        const double gs = (lap - rhs[base + m]) / c.factor;
        mine[base + m] = old + c.omega * (gs - old);
        rb = dbits(fabs(gs - old));
    }
    if (tol < 0.0) return;  // relative-residual mode: the max update is not tracked
    rb = block_max_u64(rb);
    if (threadIdx.x == 0 && rb != 0ull) sor_slot_add(ctl, it & 3, blin, rb);
}

// one wavefront
__global__ void sor_reset_kernel(SorCtl* ctl) {
    const int lane = threadIdx.x;
    for (int q = 0; q < 4; ++q)
        for (int s = lane; s < SOR_SLOTS; s += 64) ctl->slot[q][s][0] = 0ull;
    if (lane == 0) {
        ctl->converged = 0;
        ctl->iters = 0;
        ctl->last_res = 0.0;
    }
}

// one wavefront: the stop rule of iteration it (sets converged when it must
// not run); otherwise clears the slots of iteration it+1
__global__ void sor_check_kernel(SorCtl* ctl, int it, double tol) {
    const bool stop = sor_stop_check(ctl, it, tol, threadIdx.x == 0);
    if (!stop)
        for (int s = threadIdx.x; s < SOR_SLOTS; s += 64) ctl->slot[(it + 1) & 3][s][0] = 0ull;
}

// ---------------------------------------------------- multipole boundary --
// Regular solid harmonics, the same recurrence as the moment pass:
//   R_m^m = (2m-1)!! (x+iy)^m, R_{m+1}^m = (2m+1) z R_m^m,
//   (l-m) R_l^m = (2l-1) z R_{l-1}^m - (l+m-1) r^2 R_{l-2}^m.
__device__ void solid_harmonics(int lmax, double x, double y, double z,
                                double C[MP_LMAX_MAX + 1][MP_LMAX_MAX + 1],
                                double S[MP_LMAX_MAX + 1][MP_LMAX_MAX + 1]) {
    const double r2 = x * x + y * y + z * z;
    double cr = 1.0, ci = 0.0;
    for (int m = 0; m <= lmax; ++m) {
        if (m > 0) {
            const double f = (double)(2 * m - 1);
            const double nr = f * (x * cr - y * ci);
            const double ni = f * (x * ci + y * cr);
            cr = nr;
            ci = ni;
        }
        C[m][m] = cr;
        S[m][m] = ci;
        if (m + 1 <= lmax) {
            const double f = (double)(2 * m + 1);
            C[m + 1][m] = f * z * cr;
            S[m + 1][m] = f * z * ci;
        }
        for (int l = m + 2; l <= lmax; ++l) {
            const double a = (double)(2 * l - 1), b = (double)(l + m - 1), d = (double)(l - m);
            C[l][m] = (a * z * C[l - 1][m] - b * r2 * C[l - 2][m]) / d;
            S[l][m] = (a * z * S[l - 1][m] - b * r2 * S[l - 2][m]) / d;
        }
    }
}

// Multipole potential at a boundary cell centre: monopole + Cartesian
// quadrupole exactly as gravity_solver_3d.c:341-360, then the l >= 3 terms
// of the addition theorem for 1/|r - r'| (no reference counterpart).
__device__ double mp_potential(double G, int lmax, const double* mom, const double* wlm, double x,
                               double y, double z) {
    const double M = mom[0];
    const double* Q = mom + 4;
    const double rx = x - mom[1], ry = y - mom[2], rz = z - mom[3];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    double ph = 0.0;
    if (r > 0.0) {
        ph = -G * M / r;
        const double r5 = r * r * r * r * r;
        const double Qr = Q[0] * rx * rx + Q[1] * rx * ry + Q[2] * rx * rz + Q[3] * ry * rx +
                          Q[4] * ry * ry + Q[5] * ry * rz + Q[6] * rz * rx + Q[7] * rz * ry +
                          Q[8] * rz * rz;
        const double pq = -G * Qr / (2.0 * r5);
        ph += pq;
        if (lmax > 2) {
            double C[MP_LMAX_MAX + 1][MP_LMAX_MAX + 1], S[MP_LMAX_MAX + 1][MP_LMAX_MAX + 1];
            solid_harmonics(lmax, rx, ry, rz, C, S);
            const double* hi = mom + 13;
            double extra = 0.0, rp = r5;
            int q = 0;
            for (int l = 3; l <= lmax; ++l) {
                rp = rp * r * r;
                double sl = 0.0;
                for (int m = 0; m <= l; ++m) {
                    const double A = hi[q], B = hi[q + 1];
                    q += 2;
                    sl = sl + wlm[l * (MP_LMAX_MAX + 1) + m] * (A * C[l][m] + B * S[l][m]);
                }
                extra = extra + sl / rp;
            }
            ph = ph + (-G) * extra;
        }
    }
    return ph;
}

// Boundary shell: x faces (all j,k), y faces (interior i), z faces (interior
// i, j).  For an edge cell the reference's last writer is the z face, then
// the y face, and its coordinate formula differs per face
// (gravity_solver_3d.c:322-324 vs :399-401 vs :454-456).
__device__ __forceinline__ void shell_cell(long t, int nx, int ny, int nz, int& i, int& j, int& k) {
    const long nX = 2L * ny * nz, nY = 2L * (nx - 2) * nz;
    if (t < nX) {
        const long side = t / ((long)ny * nz), rem = t % ((long)ny * nz);
        i = side ? nx - 1 : 0; j = (int)(rem / nz); k = (int)(rem % nz);
    } else if (t < nX + nY) {
        const long u = t - nX, per = (long)(nx - 2) * nz;
        const long side = u / per, rem = u % per;
        i = 1 + (int)(rem / nz); j = side ? ny - 1 : 0; k = (int)(rem % nz);
    } else {
        const long u = t - nX - nY, per = (long)(nx - 2) * (ny - 2);
        const long side = u / per, rem = u % per;
        i = 1 + (int)(rem / (ny - 2)); j = 1 + (int)(rem % (ny - 2)); k = side ? nz - 1 : 0;
    }
}

__global__ void __launch_bounds__(256) sor_boundary_kernel(GridDims g, int bc, double G, int lmax,
                                                           const double* __restrict__ mom,
                                                           const double* __restrict__ wlm,
                                                           double* __restrict__ phi_r,
                                                           double* __restrict__ phi_b) {
    const long tot = 2L * g.ny * g.nz + 2L * (g.nx - 2) * g.nz + 2L * (g.nx - 2) * (g.ny - 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    int i, j, k;
    shell_cell(t, g.nx, g.ny, g.nz, i, j, k);
    double v;
    if (bc == 1) {  // isolated: multipole expansion
        double x, y, z;
        if (k == 0 || k == g.nz - 1) {
            x = g.xmin + (i + 0.5) * g.dx;
            y = g.ymin + (j + 0.5) * g.dy;
            z = (k == 0) ? g.zmin + 0.5 * g.dz : g.zmax - 0.5 * g.dz;
        } else if (j == 0 || j == g.ny - 1) {
            x = g.xmin + (i + 0.5) * g.dx;
            y = (j == 0) ? g.ymin + 0.5 * g.dy : g.ymax - 0.5 * g.dy;
            z = g.zmin + (k + 0.5) * g.dz;
        } else {
            x = (i == 0) ? g.xmin + 0.5 * g.dx : g.xmax - 0.5 * g.dx;
            y = g.ymin + (j + 0.5) * g.dy;
            z = g.zmin + (k + 0.5) * g.dz;
        }
        v = mp_potential(G, lmax, mom, wlm, x, y, z);
    } else if (bc == 0) {  // periodic copy, gravity_solver_3d.c:273-299
        const int si = (i == 0) ? g.nx - 1 : i;
        const int sj = (j == 0) ? g.ny - 1 : j;
        const int sk = (k == 0) ? g.nz - 1 : k;
        v = phi_at(g, phi_r, phi_b, si, sj, sk);
    } else {
        return;  // Dirichlet: boundary left as it is
    }
    (((i + j + k) & 1) ? phi_b : phi_r)[sidx(g, i, j, k)] = v;
}

// ------------------------------------------------- the 2D branch (nz = 1) --
// gravity_solver_3d.c:569-634: one z level, no z derivatives, the potential
// in the natural layout phi[i ny + j]; the stop rule, slots and check kernel
// are the 3D ones.  One thread per cell of the colour: row i = 1 + blockIdx.y,
// j = j0 + 2 m with (i + j) % 2 == colour.
__global__ void __launch_bounds__(256) sor2d_sweep_kernel(GridDims g, SorCoef c, int colour,
                                                          int it, double tol, double four_pi_G,
                                                          double* __restrict__ phi,
                                                          const double* __restrict__ rho,
                                                          SorCtl* ctl) {
    if (tol >= 0.0 && ctl->converged) return;
    const unsigned blin = blockIdx.y * gridDim.x + blockIdx.x;
    const int i = 1 + blockIdx.y;
    const int j = 1 + ((i + 1 + colour) & 1) + 2 * (int)(blockIdx.x * 256 + threadIdx.x);
    unsigned long long rb = 0ull;
    if (j <= g.ny - 2) {
        const long idx = (long)i * g.ny + j;
        const double old = phi[idx];
        const double lap = (phi[idx - g.ny] + phi[idx + g.ny]) * c.dx2i +
                           (phi[idx - 1] + phi[idx + 1]) * c.dy2i;
        const double rhs = four_pi_G * rho[idx];
        const double gs = (lap - rhs) / c.factor;
        phi[idx] = old + c.omega * (gs - old);
        rb = dbits(fabs(gs - old));
    }
    if (tol < 0.0) return;
    rb = block_max_u64(rb);
    if (threadIdx.x == 0 && rb != 0ull) sor_slot_add(ctl, it & 3, blin, rb);
}

// Boundary of the 2D branch: the x faces (all j) and the y faces (all i),
// no z faces (gravity_solver_3d.c:318-400, :450).  The y face writes the
// corners last, with its coordinate formula.
__global__ void __launch_bounds__(256) sor2d_boundary_kernel(GridDims g, int bc, double G,
                                                             int lmax,
                                                             const double* __restrict__ mom,
                                                             const double* __restrict__ wlm,
                                                             double* __restrict__ phi) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2L * g.ny + 2L * (g.nx - 2)) return;
    int i, j;
    if (t < 2L * g.ny) {
        i = t < g.ny ? 0 : g.nx - 1;
        j = (int)(t % g.ny);
    } else {
        const long u = t - 2L * g.ny;
        i = 1 + (int)(u % (g.nx - 2));
        j = u < g.nx - 2 ? 0 : g.ny - 1;
    }
    double v;
    if (bc == 1) {
        const double z = g.zmin + (0 + 0.5) * g.dz;
        double x, y;
        if (j == 0 || j == g.ny - 1) {
            x = g.xmin + (i + 0.5) * g.dx;
            y = (j == 0) ? g.ymin + 0.5 * g.dy : g.ymax - 0.5 * g.dy;
        } else {
            x = (i == 0) ? g.xmin + 0.5 * g.dx : g.xmax - 0.5 * g.dx;
            y = g.ymin + (j + 0.5) * g.dy;
        }
        v = mp_potential(G, lmax, mom, wlm, x, y, z);
    } else if (bc == 0) {  // periodic copies: x then y, gravity_solver_3d.c:273-289
        const int si = (i == 0) ? g.nx - 1 : i;
        const int sj = (j == 0) ? g.ny - 1 : j;
        v = phi[(long)si * g.ny + sj];
    } else {
        return;
    }
    phi[(long)i * g.ny + j] = v;
}

// g = -grad phi of the 2D branch (gravity_solver_3d.c:764-809): interior
// (i, j) from centred differences, gz = 0, the boundary ring all zero.
__global__ void __launch_bounds__(256) accel2d_kernel(GridDims g, const double* __restrict__ phi,
                                                      double* __restrict__ gx,
                                                      double* __restrict__ gy,
                                                      double* __restrict__ gz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.nx * g.ny) return;
    const int i = (int)(t / g.ny), j = (int)(t % g.ny);
    double ax = 0.0, ay = 0.0;
    if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2) {
        ax = -(phi[t + g.ny] - phi[t - g.ny]) / (2.0 * g.dx);
        ay = -(phi[t + 1] - phi[t - 1]) / (2.0 * g.dy);
    }
    gx[t] = ax;
    gy[t] = ay;
    gz[t] = 0.0;
}

// --------------------------------------------------------- acceleration --
// gravity_solver_3d.c:820-897.  The sequential x, y, z boundary copies leave
// every boundary cell with the interior value at the clamped index.
// One thread per cell: k along the 64 lanes of a row, 4 rows (j) per block,
// one plane (i) per blockIdx.z, so no index needs an integer division.
__global__ void __launch_bounds__(256) acceleration_kernel(GridDims g, const double* __restrict__ pr,
                                                           const double* __restrict__ pb,
                                                           double* __restrict__ gx,
                                                           double* __restrict__ gy,
                                                           double* __restrict__ gz) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long t = ((long)i * g.ny + j) * g.nz + k;
    const int a = min(max(i, 1), g.nx - 2);
    const int b = min(max(j, 1), g.ny - 2);
    const int c = min(max(k, 1), g.nz - 2);
    gx[t] = -(phi_at(g, pr, pb, a + 1, b, c) - phi_at(g, pr, pb, a - 1, b, c)) / (2.0 * g.dx);
    gy[t] = -(phi_at(g, pr, pb, a, b + 1, c) - phi_at(g, pr, pb, a, b - 1, c)) / (2.0 * g.dy);
    gz[t] = -(phi_at(g, pr, pb, a, b, c + 1) - phi_at(g, pr, pb, a, b, c - 1)) / (2.0 * g.dz);
}

// -------------------------------------------------------------- moments --
// Each quantity is summed as a pairwise double-double tree over the flat
// cell index (zero padded to a power of two): a block reduces 256
// consecutive cells (lane butterfly, then the four wavefronts pairwise), and
// reduce passes fold 256 consecutive partials at a time.  The result does
// not depend on the launch shape beyond that tree, and it is accurate to a
// few units in the 106th bit before the final rounding.

__device__ __forceinline__ void cell_coord(const GridDims& g, long t, int& i, int& j, int& k) {
    k = (int)(t % g.nz);
    const long ij = t / g.nz;
    j = (int)(ij % g.ny);
    i = (int)(ij / g.ny);
}

constexpr int MQC = 16;         // quantities per LDS round
constexpr int MROW = 256 + 16;  // one padding entry per 16 cells: the 16-leaf
                                // subtrees of 16 threads fall on distinct banks

__device__ __forceinline__ int mslot(int c) { return c + (c >> 4); }

// The pairwise double-double tree over the block's 256 consecutive cells of
// each of NQ terms: 16 threads per quantity each sum a 16-leaf subtree, then
// the 16 subtree sums meet pairwise across those 16 lanes (partners 1, 2, 4,
// 8 apart); the block's sum of quantity q goes to part[q][blockIdx.x].  Two
// leaves meet in an exact two-sum (their double-double sum with zero low
// parts), every later node in a double-double addition, the left operand
// always the lower index.
template <int NQ>
__device__ __forceinline__ void block_tree_terms(const double (&term)[NQ], double* __restrict__ part,
                                                 double (*s_term)[MROW]) {
    const int tid = threadIdx.x;
    const int qq = tid >> 4, st = tid & 15;
#pragma unroll
    for (int c0 = 0; c0 < NQ; c0 += MQC) {
        if (c0 > 0) __syncthreads();  // the previous round's reads of s_term are done
#pragma unroll
        for (int k = 0; k < MQC; ++k)
            if (c0 + k < NQ) s_term[k][mslot(tid)] = term[c0 + k];
        __syncthreads();
        if (c0 + qq < NQ) {  // whole groups of 16 lanes
            const double* leaf = &s_term[qq][mslot(st * 16)];
            DD v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = two_sum(leaf[2 * e], leaf[2 * e + 1]);
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = dd_add(v[2 * e], v[2 * e + 1]);
            v[0] = dd_add(v[0], v[1]);
            v[1] = dd_add(v[2], v[3]);
            DD sub = dd_add(v[0], v[1]);
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                DD w;
                w.hi = __shfl_xor(sub.hi, o, 64);
                w.lo = __shfl_xor(sub.lo, o, 64);
                if ((st & (2 * o - 1)) == 0) sub = dd_add(sub, w);
            }
            if (st == 0) {
                const long o = 2 * ((long)(c0 + qq) * gridDim.x + blockIdx.x);
                part[o] = sub.hi;
                part[o + 1] = sub.lo;
            }
        }
    }
}

// Where a block's terms go: the double-double tree (part = block partials),
// or, for the reference's sequential order, stored as they are (part =
// terms[q][npad], cell index fastest; the grid covers npad cells and the
// padding cells store zeros).
template <int NQ>
__device__ __forceinline__ void emit_terms(bool store, const double (&term)[NQ], long t, long npad,
                                           double* __restrict__ part, double (*s_term)[MROW]) {
    if (store) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[(long)q * npad + t] = term[q];
    } else {
        block_tree_terms<NQ>(term, part, s_term);
    }
}

// pass 1: M and the first moments (gravity_solver_3d.c:182-198)
template <bool STORE>
__global__ void __launch_bounds__(256) moments_pass1_kernel(GridDims g,
                                                            const double* __restrict__ rho,
                                                            double* __restrict__ part, long npad) {
    __shared__ double s_term[MQC][MROW];
    const long n = (long)g.nx * g.ny * g.nz;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    double term[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < n) {
        const double r = rho[t];
        if (r > 0.0) {
            int i, j, k;
            cell_coord(g, t, i, j, k);
            const double dV = g.dx * g.dy * g.dz;
            const double x = g.xmin + (i + 0.5) * g.dx;
            const double y = g.ymin + (j + 0.5) * g.dy;
            const double z = g.zmin + (k + 0.5) * g.dz;
            const double dm = r * dV;
            term[0] = dm; term[1] = dm * x; term[2] = dm * y; term[3] = dm * z;
        }
    }
    emit_terms<4>(STORE, term, t, npad, part, s_term);
}

constexpr int mp_nhi_c(int lmax) {
    int n = 0;
    for (int l = 3; l <= lmax; ++l) n += 2 * (l + 1);
    return n;
}

// pass 2: Q_ij (gravity_solver_3d.c:226-258) and A_lm, B_lm for l >= 3
template <int LMAX, bool STORE>
__global__ void __launch_bounds__(256) moments_pass2_kernel(GridDims g,
                                                            const double* __restrict__ rho,
                                                            const double* __restrict__ mom,
                                                            double* __restrict__ part, long npad) {
    constexpr int NQ = 9 + mp_nhi_c(LMAX);
    __shared__ double s_term[MQC][MROW];
    const long n = (long)g.nx * g.ny * g.nz;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const double r = (t < n) ? rho[t] : 0.0;
    const bool in = (t < n) && (r > 0.0);
    double x = 0.0, y = 0.0, z = 0.0, dm = 0.0, r2 = 0.0;
    if (in) {
        int i, j, k;
        cell_coord(g, t, i, j, k);
        const double dV = g.dx * g.dy * g.dz;
        x = g.xmin + (i + 0.5) * g.dx - mom[1];
        y = g.ymin + (j + 0.5) * g.dy - mom[2];
        z = g.zmin + (k + 0.5) * g.dz - mom[3];
        dm = r * dV;
        r2 = x * x + y * y + z * z;
    }
    double term[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) term[q] = 0.0;
    if (in) {
        term[0] = dm * (3.0 * x * x - r2);
        term[1] = dm * (3.0 * x * y);
        term[2] = dm * (3.0 * x * z);
        term[3] = dm * (3.0 * y * x);
        term[4] = dm * (3.0 * y * y - r2);
        term[5] = dm * (3.0 * y * z);
        term[6] = dm * (3.0 * z * x);
        term[7] = dm * (3.0 * z * y);
        term[8] = dm * (3.0 * z * z - r2);
    }
    if (LMAX >= 3) {
        // walk the recurrence per m; emit each (l, m) with l >= 3 at its slot
        double cr = 1.0, ci = 0.0;
#pragma unroll
        for (int m = 0; m <= LMAX; ++m) {
            if (m > 0) {
                const double f = (double)(2 * m - 1);
                const double nr = f * (x * cr - y * ci);
                const double ni = f * (x * ci + y * cr);
                cr = nr;
                ci = ni;
            }
            double c2 = cr, s2 = ci;  // R_{l-2}, R_{l-1} as the walk advances
            double c1 = 0.0, s1 = 0.0;
#pragma unroll
            for (int l = m; l <= LMAX; ++l) {
                double cl, sl;
                if (l == m) {
                    cl = cr; sl = ci;
                } else if (l == m + 1) {
                    const double f = (double)(2 * m + 1);
                    cl = f * z * cr; sl = f * z * ci;
                } else {
                    const double a = (double)(2 * l - 1), b = (double)(l + m - 1), d = (double)(l - m);
                    cl = div_by_const(a * z * c1 - b * r2 * c2, d);
                    sl = div_by_const(a * z * s1 - b * r2 * s2, d);
                }
                if (l == m) {
                    c1 = cl; s1 = sl;
                } else {
                    c2 = c1; s2 = s1; c1 = cl; s1 = sl;
                }
                if (l >= 3) {
                    const int q = 9 + mp_nhi_c(l - 1) + 2 * m;
                    term[q] = in ? dm * cl : 0.0;
                    term[q + 1] = in ? dm * sl : 0.0;
                }
            }
        }
    }
    emit_terms<NQ>(STORE, term, t, npad, part, s_term);
}

// Fold 256 consecutive partials of quantity q (zero padded) pairwise.
__global__ void __launch_bounds__(256) moments_reduce_kernel(const double* __restrict__ in,
                                                             long nin, double* __restrict__ out,
                                                             long nout) {
    __shared__ DD s[256];
    const int q = blockIdx.y;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    DD v{0.0, 0.0};
    if (idx < nin) {
        v.hi = in[2 * ((long)q * nin + idx)];
        v.lo = in[2 * ((long)q * nin + idx) + 1];
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        DD u{0.0, 0.0};
        if (threadIdx.x < h) u = dd_add(s[2 * threadIdx.x], s[2 * threadIdx.x + 1]);
        __syncthreads();
        if (threadIdx.x < h) s[threadIdx.x] = u;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * ((long)q * nout + blockIdx.x)] = s[0].hi;
        out[2 * ((long)q * nout + blockIdx.x) + 1] = s[0].lo;
    }
}

// Round the sums to double; pass 1 also forms the centre of mass
// (gravity_solver_3d.c:200-209).
__global__ void moments_finish_kernel(const double* __restrict__ sums, int nq, int pass,
                                      GridDims g, double* __restrict__ mom) {
    if (threadIdx.x != 0) return;
    if (pass == 1) {
        const double M = sums[0];
        mom[0] = M;
        if (M > 0.0) {
            mom[1] = sums[2] / M;
            mom[2] = sums[4] / M;
            mom[3] = sums[6] / M;
        } else {
            mom[1] = 0.5 * (g.xmin + g.xmax);
            mom[2] = 0.5 * (g.ymin + g.ymax);
            mom[3] = 0.5 * (g.zmin + g.zmax);
        }
    } else {
        for (int q = 0; q < nq; ++q) mom[4 + q] = sums[2 * q];
    }
}

// ---------------------------------------------------- defect correction --
// r = f - L phi on interior cells (0 on the boundary), both colours of
// split entry m of row (i, j); the block's sums of r^2 and f^2 go to
// part[block] and residual_finish_kernel adds the blocks in a fixed order
// (no atomics: the result is reproducible).  A block marches RES_PLANES
// consecutive x planes.  Each thread keeps the centre values of planes
// i - 1 .. i + 1 of both colours of its (j, m) in registers (the x
// neighbours, the centre, and the z neighbour at the same split index), and
// the loads of plane i + 1's other operands and plane i + 2's centres are
// issued before plane i is computed, so one plane of loads is in flight.
#ifndef HYD_RES_PLANES
#define HYD_RES_PLANES 8  // 16 and 4 measured 4% slower at 256^3 (kernel + final sum)
#endif
constexpr int RES_PLANES = HYD_RES_PLANES;

struct ResOperands {  // one colour of one cell: f, y neighbours, the other z neighbour
    double f, ym, yp, zx;
};

__device__ __forceinline__ ResOperands res_load(const GridDims& g, int i, int j, int m, int colour,
                                                const double* __restrict__ pr,
                                                const double* __restrict__ pb,
                                                const double* __restrict__ fr,
                                                const double* __restrict__ fb) {
    ResOperands o{0.0, 0.0, 0.0, 0.0};
    const int p = (i + j + colour) & 1, k = 2 * m + p;
    if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2 && k >= 1 && k <= g.nz - 2) {
        const long base = ((long)i * g.ny + j) * g.nzh;
        const double* other = colour ? pr : pb;
        o.f = (colour ? fb : fr)[base + m];
        o.ym = other[base - g.nzh + m];
        o.yp = other[base + g.nzh + m];
        o.zx = other[base + (p ? m + 1 : m - 1)];  // the z pair is this and other[m]
    }
    return o;
}

__global__ void __launch_bounds__(256) residual_kernel(GridDims g, SorCoef c,
                                                       const double* __restrict__ pr,
                                                       const double* __restrict__ pb,
                                                       const double* __restrict__ fr,
                                                       const double* __restrict__ fb,
                                                       float* __restrict__ rr,
                                                       float* __restrict__ rbk,
                                                       double* __restrict__ part) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * SOR_ROWS + (threadIdx.x >> 6);
    const int i0 = blockIdx.z * RES_PLANES;
    const int i1 = i0 + RES_PLANES < g.nx ? i0 + RES_PLANES : g.nx;
    double s_rr = 0.0, s_ff = 0.0;
    if (m < g.nzh && j < g.ny) {
        const long sxh = (long)g.ny * g.nzh;
        const long row = (long)j * g.nzh + m;
        double r_prev = 0.0, b_prev = 0.0, r_next = 0.0, b_next = 0.0;
        if (i0 > 0) {
            r_prev = pr[(long)(i0 - 1) * sxh + row];
            b_prev = pb[(long)(i0 - 1) * sxh + row];
        }
        double r_cur = pr[(long)i0 * sxh + row], b_cur = pb[(long)i0 * sxh + row];
        if (i0 + 1 < g.nx) {
            r_next = pr[(long)(i0 + 1) * sxh + row];
            b_next = pb[(long)(i0 + 1) * sxh + row];
        }
        ResOperands o0 = res_load(g, i0, j, m, 0, pr, pb, fr, fb);
        ResOperands o1 = res_load(g, i0, j, m, 1, pr, pb, fr, fb);
        for (int i = i0; i < i1; ++i) {
            // next iteration's loads, in flight while plane i is computed
            double r_nn = 0.0, b_nn = 0.0;
            if (i + 2 < g.nx) {
                r_nn = pr[(long)(i + 2) * sxh + row];
                b_nn = pb[(long)(i + 2) * sxh + row];
            }
            ResOperands n0{0.0, 0.0, 0.0, 0.0}, n1{0.0, 0.0, 0.0, 0.0};
            if (i + 1 < i1) {
                n0 = res_load(g, i + 1, j, m, 0, pr, pb, fr, fb);
                n1 = res_load(g, i + 1, j, m, 1, pr, pb, fr, fb);
            }
            const long base = (long)i * sxh + (long)j * g.nzh;
#pragma unroll
            for (int colour = 0; colour < 2; ++colour) {
                const int k = 2 * m + ((i + j + colour) & 1);
                if (k >= g.nz) continue;
                float* rout = colour ? rbk : rr;
                if (i < 1 || i > g.nx - 2 || j < 1 || j > g.ny - 2 || k < 1 || k > g.nz - 2) {
                    if (rout) rout[base + m] = 0.0f;
                    continue;
                }
                const ResOperands& o = colour ? o1 : o0;
                const double lap = ((colour ? r_prev : b_prev) + (colour ? r_next : b_next)) *
                                       c.dx2i +
                                   (o.ym + o.yp) * c.dy2i +
                                   (o.zx + (colour ? r_cur : b_cur)) * c.dz2i -
                                   c.factor * (colour ? b_cur : r_cur);
                const double r = o.f - lap;
                if (rout) rout[base + m] = (float)r;
                s_rr += r * r;
                s_ff += o.f * o.f;
            }
            r_prev = r_cur;
            b_prev = b_cur;
            r_cur = r_next;
            b_cur = b_next;
            r_next = r_nn;
            b_next = b_nn;
            o0 = n0;
            o1 = n1;
        }
    }
    __shared__ double sp[2][4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s_rr += __shfl_xor(s_rr, o, 64);
        s_ff += __shfl_xor(s_ff, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sp[0][threadIdx.x >> 6] = s_rr;
        sp[1][threadIdx.x >> 6] = s_ff;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const long b = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part[2 * b] = (sp[0][0] + sp[0][1]) + (sp[0][2] + sp[0][3]);
        part[2 * b + 1] = (sp[1][0] + sp[1][1]) + (sp[1][2] + sp[1][3]);
    }
}

__global__ void __launch_bounds__(256) residual_finish_kernel(const double* __restrict__ part,
                                                              long nblk, double* acc) {
    double a = 0.0, b = 0.0;
    for (long q = threadIdx.x; q < nblk; q += 256) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    __shared__ double s[2][256];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] = s[0][0];
        acc[1] = s[1][0];
    }
}

// Sum over the interior cells of (r - L e)^2: the residual of the start
// phi + e when r is phi's residual (L e in FP32 as the inner sweeps apply
// it, squares and sums in FP64), per block into part[2 b] (part[2 b + 1] =
// 0) in the geometry of residual_kernel, so residual_finish_kernel adds the
// blocks in the same fixed order.
__global__ void __launch_bounds__(256) start_residual_sq_kernel(GridDims g, float dx2i,
                                                                float dy2i, float dz2i,
                                                                float factor,
                                                                const float* __restrict__ er,
                                                                const float* __restrict__ eb,
                                                                const float* __restrict__ rr,
                                                                const float* __restrict__ rb,
                                                                double* __restrict__ part) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * SOR_ROWS + (threadIdx.x >> 6);
    const int i0 = blockIdx.z * RES_PLANES;
    const int i1 = i0 + RES_PLANES < g.nx ? i0 + RES_PLANES : g.nx;
    double s = 0.0;
    if (m < g.nzh && j >= 1 && j <= g.ny - 2) {
        const long sxh = (long)g.ny * g.nzh;
        for (int i = i0 > 1 ? i0 : 1; i < i1 && i <= g.nx - 2; ++i) {
            const long base = (long)i * sxh + (long)j * g.nzh;
#pragma unroll
            for (int colour = 0; colour < 2; ++colour) {
                const int k = 2 * m + ((i + j + colour) & 1);
                if (k < 1 || k > g.nz - 2) continue;
                const float* mine = colour ? eb : er;
                const float* other = colour ? er : eb;
                const float lap = (other[base - sxh + m] + other[base + sxh + m]) * dx2i +
                                  (other[base - g.nzh + m] + other[base + g.nzh + m]) * dy2i +
                                  (other[base + ((k - 1) >> 1)] + other[base + ((k + 1) >> 1)]) *
                                      dz2i -
                                  factor * mine[base + m];
                const double q = (double)((colour ? rb : rr)[base + m] - lap);
                s += q * q;
            }
        }
    }
    __shared__ double sp[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sp[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long b = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part[2 * b] = (sp[0] + sp[1]) + (sp[2] + sp[3]);
        part[2 * b + 1] = 0.0;
    }
}

// One colour of red-black SOR on L e = r in FP32 (same indexing as
// sor_sweep_kernel); gs = (lap - r) * (1 / factor).
__global__ void __launch_bounds__(256) sor_sweep_f32_kernel(GridDims g, float dx2i, float dy2i,
                                                            float dz2i, float inv_factor,
                                                            float omega, int colour,
                                                            float* __restrict__ mine,
                                                            const float* __restrict__ other,
                                                            const float* __restrict__ rhs) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * SOR_ROWS + (threadIdx.x >> 6);
    const int i = 1 + blockIdx.z;
    const int k = 2 * m + ((i + j + colour) & 1);
    if (j > g.ny - 2 || k < 1 || k > g.nz - 2) return;
    const long base = ((long)i * g.ny + j) * g.nzh;
    const long sxh = (long)g.ny * g.nzh;
    const float lap = (other[base - sxh + m] + other[base + sxh + m]) * dx2i +
                      (other[base - g.nzh + m] + other[base + g.nzh + m]) * dy2i +
                      (other[base + ((k - 1) >> 1)] + other[base + ((k + 1) >> 1)]) * dz2i;
    const float old = mine[base + m];
    const float gs = (lap - rhs[base + m]) * inv_factor;
    mine[base + m] = old + omega * (gs - old);
}

// The same sweep, four split entries per thread with 16-byte loads (rows
// whose entry count is a multiple of 4).  Cells off the interior keep their
// value (the store writes it back unchanged).  Row r of the block's 256
// threads: groups of 4 entries along the row, then rows, then planes.
__global__ void __launch_bounds__(256) sor_sweep_f32x4_kernel(GridDims g, float dx2i, float dy2i,
                                                              float dz2i, float inv_factor,
                                                              float omega, int colour,
                                                              float* __restrict__ mine,
                                                              const float* __restrict__ other,
                                                              const float* __restrict__ rhs,
                                                              unsigned groups_per_row,
                                                              unsigned total) {
    const unsigned t = blockIdx.x * 256u + threadIdx.x;
    if (t >= total) return;
    const unsigned row = t / groups_per_row, q = t - row * groups_per_row;
    const int j = 1 + (int)(row % (unsigned)(g.ny - 2));
    const int i = 1 + (int)(row / (unsigned)(g.ny - 2));
    const int m0 = 4 * (int)q;
    const int par = (i + j + colour) & 1;  // cell k = 2m + par
    const long base = ((long)i * g.ny + j) * g.nzh + m0;
    const long sxh = (long)g.ny * g.nzh;
    const float4 xo = *reinterpret_cast<const float4*>(mine + base);
    const float4 xr = *reinterpret_cast<const float4*>(rhs + base);
    const float4 im = *reinterpret_cast<const float4*>(other + base - sxh);
    const float4 ip = *reinterpret_cast<const float4*>(other + base + sxh);
    const float4 jm = *reinterpret_cast<const float4*>(other + base - g.nzh);
    const float4 jp = *reinterpret_cast<const float4*>(other + base + g.nzh);
    const float4 kc = *reinterpret_cast<const float4*>(other + base);
    // the entry beyond the group on the side the z neighbours reach
    float edge = 0.0f;
    if (par == 0 && m0 > 0) edge = other[base - 1];
    if (par == 1 && m0 + 4 < g.nzh) edge = other[base + 4];
    const float o[6] = {par ? kc.x : edge, kc.x, kc.y, kc.z, kc.w, par ? edge : kc.w};
    // cell e: k - 1 -> o[e + par], k + 1 -> o[e + par + 1] (o[0] = entry m0 - 1)
    const float old[4] = {xo.x, xo.y, xo.z, xo.w};
    const float rr[4] = {xr.x, xr.y, xr.z, xr.w};
    const float vim[4] = {im.x, im.y, im.z, im.w}, vip[4] = {ip.x, ip.y, ip.z, ip.w};
    const float vjm[4] = {jm.x, jm.y, jm.z, jm.w}, vjp[4] = {jp.x, jp.y, jp.z, jp.w};
    float out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = 2 * (m0 + e) + par;
        const float okm = par ? o[e + 1] : o[e];
        const float okp = par ? o[e + 2] : o[e + 1];
        const float lap = (vim[e] + vip[e]) * dx2i + (vjm[e] + vjp[e]) * dy2i + (okm + okp) * dz2i;
        const float gs = (lap - rr[e]) * inv_factor;
        const float nv = old[e] + omega * (gs - old[e]);
        out[e] = (k >= 1 && k <= g.nz - 2) ? nv : old[e];
    }
    *reinterpret_cast<float4*>(mine + base) = make_float4(out[0], out[1], out[2], out[3]);
}

__global__ void correct_kernel(GridDims g, double* __restrict__ pr, double* __restrict__ pb,
                               float* __restrict__ er, float* __restrict__ eb) {
    const long n = (long)g.nx * g.ny * g.nzh;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        pr[t] = pr[t] + (double)er[t];
        pb[t] = pb[t] + (double)eb[t];
        er[t] = 0.0f;
        eb[t] = 0.0f;
    }
}

// ----------------------------------------------------------- launchers --
static unsigned grid_for(long n, long cap = 16384) {
    long nb = (n + 255) / 256;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

static dim3 sor_grid(const GridDims& g) {
    return dim3((unsigned)((g.nzh + 63) / 64), (unsigned)((g.ny - 2 + SOR_ROWS - 1) / SOR_ROWS),
                (unsigned)(g.nx - 2));
}

long moments_scratch_len(long ncells, int lmax) {
    const long nb = (ncells + 255) / 256;
    const long nq = mp_n_pass2(lmax) > 4 ? mp_n_pass2(lmax) : 4;
    const long nb2 = (nb + 255) / 256;
    return 2 * nq * (nb + nb2 + 2) + 2 * nq;
}

// Reduce `nb` partial double-doubles per quantity down to one, in place of
// the scratch area; returns the pointer holding nq final pairs.
static double* reduce_partials(double* part, long nb, int nq, hipStream_t st) {
    double* cur = part;
    double* nxt = part + 2L * nq * nb;
    long n = nb;
    while (n > 1) {
        const long nout = (n + 255) / 256;
        hipLaunchKernelGGL(moments_reduce_kernel, dim3((unsigned)nout, nq), dim3(256), 0, st, cur,
                           n, nxt, nout);
        double* tmp = cur;
        cur = nxt;
        nxt = tmp;
        n = nout;
    }
    return cur;
}

template <bool STORE>
static hipError_t launch_pass2(const GridDims& g, const double* rho, int lmax, const double* mom,
                               double* part, long npad, long nb, hipStream_t st) {
    switch (lmax) {
#define MP_CASE(L)                                                                          \
    case L:                                                                                 \
        hipLaunchKernelGGL((moments_pass2_kernel<L, STORE>), dim3((unsigned)nb), dim3(256), 0, \
                           st, g, rho, mom, part, npad);                                   \
        break;
        MP_CASE(2) MP_CASE(3) MP_CASE(4) MP_CASE(5) MP_CASE(6) MP_CASE(7) MP_CASE(8)
#undef MP_CASE
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_moments(const GridDims& g, const double* rho, int lmax, double* scratch,
                          double* mom_out, hipStream_t st) {
    const long n = (long)g.nx * g.ny * g.nz;
    const long nb = (n + 255) / 256;
    hipLaunchKernelGGL(moments_pass1_kernel<false>, dim3((unsigned)nb), dim3(256), 0, st, g, rho,
                       scratch, 0L);
    double* s1 = reduce_partials(scratch, nb, 4, st);
    hipLaunchKernelGGL(moments_finish_kernel, dim3(1), dim3(64), 0, st, s1, 4, 1, g, mom_out);
    const int nq = mp_n_pass2(lmax);
    const hipError_t e = launch_pass2<false>(g, rho, lmax, mom_out, scratch, 0L, nb, st);
    if (e != hipSuccess) return e;
    double* s2 = reduce_partials(scratch, nb, nq, st);
    hipLaunchKernelGGL(moments_finish_kernel, dim3(1), dim3(64), 0, st, s2, nq, 2, g, mom_out);
    return hipGetLastError();
}

long moments_seq_npad(long ncells) { return seqsum_npad(ncells); }

static int seq_nq(int lmax) { return mp_n_pass2(lmax) > 4 ? mp_n_pass2(lmax) : 4; }

long moments_seq_len(long ncells, int lmax) {
    const long npad = seqsum_npad(ncells);
    const int nq = seq_nq(lmax);
    return nq * npad + seqsum_workspace_len(npad, nq) + 2L * nq;
}

hipError_t launch_moments_sequential(const GridDims& g, const double* rho, int lmax, int method,
                                     double* buf, double* mom_out, hipStream_t st) {
    const long n = (long)g.nx * g.ny * g.nz;
    const long npad = seqsum_npad(n);
    const long nb = npad / 256;
    const int nq = mp_n_pass2(lmax);
    double* terms = buf;
    double* ws = terms + (long)seq_nq(lmax) * npad;
    double* sums = ws + seqsum_workspace_len(npad, seq_nq(lmax));
    hipLaunchKernelGGL(moments_pass1_kernel<true>, dim3((unsigned)nb), dim3(256), 0, st, g, rho,
                       terms, npad);
    hipError_t e = launch_seqsum(method, terms, npad, 4, ws, sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(moments_finish_kernel, dim3(1), dim3(64), 0, st, sums, 4, 1, g, mom_out);
    e = launch_pass2<true>(g, rho, lmax, mom_out, terms, npad, nb, st);
    if (e != hipSuccess) return e;
    e = launch_seqsum(method, terms, npad, nq, ws, sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(moments_finish_kernel, dim3(1), dim3(64), 0, st, sums, nq, 2, g, mom_out);
    return hipGetLastError();
}

hipError_t launch_natural_to_split(const GridDims& g, const double* nat, double scale,
                                   double* red, double* black, hipStream_t st) {
    const long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(natural_to_split_kernel, dim3(grid_for(n)), dim3(256), 0, st, g, nat, scale,
                       red, black);
    return hipGetLastError();
}

hipError_t launch_split_to_natural(const GridDims& g, const double* red, const double* black,
                                   double* nat, hipStream_t st) {
    const long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(split_to_natural_kernel, dim3(grid_for(n)), dim3(256), 0, st, g, red, black,
                       nat);
    return hipGetLastError();
}

hipError_t launch_sor_sweep(const GridDims& g, const SorCoef& c, int colour, int it, double tol,
                            double* phi_r, double* phi_b, const double* rhs_r,
                            const double* rhs_b, SorCtl* ctl, hipStream_t st) {
    double* mine = colour ? phi_b : phi_r;
    const double* other = colour ? phi_r : phi_b;
    const double* rhs = colour ? rhs_b : rhs_r;
    hipLaunchKernelGGL(sor_sweep_kernel, sor_grid(g), dim3(256), 0, st, g, c, colour, it, tol,
                       mine, other, rhs, ctl);
    return hipGetLastError();
}

hipError_t launch_sor_check(SorCtl* ctl, int it, double tol, hipStream_t st) {
    hipLaunchKernelGGL(sor_check_kernel, dim3(1), dim3(64), 0, st, ctl, it, tol);
    return hipGetLastError();
}

hipError_t launch_sor_boundary(const GridDims& g, int bc, double G, int lmax, const double* mom,
                               const double* wlm, double* phi_r, double* phi_b, hipStream_t st) {
    if (bc == 2) return hipSuccess;
    const long tot = 2L * g.ny * g.nz + 2L * (g.nx - 2) * g.nz + 2L * (g.nx - 2) * (g.ny - 2);
    hipLaunchKernelGGL(sor_boundary_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       g, bc, G, lmax, mom, wlm, phi_r, phi_b);
    return hipGetLastError();
}

hipError_t launch_sor2d_sweep(const GridDims& g, const SorCoef& c, int colour, int it, double tol,
                              double four_pi_G, double* phi, const double* rho, SorCtl* ctl,
                              hipStream_t st) {
    const int per_row = (g.ny - 1) / 2;  // cells of one colour in a row, at most
    const dim3 grid((unsigned)((per_row + 255) / 256), (unsigned)(g.nx - 2));
    hipLaunchKernelGGL(sor2d_sweep_kernel, grid, dim3(256), 0, st, g, c, colour, it, tol,
                       four_pi_G, phi, rho, ctl);
    return hipGetLastError();
}

hipError_t launch_sor2d_boundary(const GridDims& g, int bc, double G, int lmax, const double* mom,
                                 const double* wlm, double* phi, hipStream_t st) {
    if (bc == 2) return hipSuccess;
    const long tot = 2L * g.ny + 2L * (g.nx - 2);
    hipLaunchKernelGGL(sor2d_boundary_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       st, g, bc, G, lmax, mom, wlm, phi);
    return hipGetLastError();
}

hipError_t launch_accel2d(const GridDims& g, const double* phi, double* gx, double* gy, double* gz,
                          hipStream_t st) {
    const long n = (long)g.nx * g.ny;
    hipLaunchKernelGGL(accel2d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, phi,
                       gx, gy, gz);
    return hipGetLastError();
}

hipError_t launch_sor_reset(SorCtl* ctl, hipStream_t st) {
    hipLaunchKernelGGL(sor_reset_kernel, dim3(1), dim3(64), 0, st, ctl);
    return hipGetLastError();
}

hipError_t launch_acceleration(const GridDims& g, const double* phi_r, const double* phi_b,
                               double* gx, double* gy, double* gz, hipStream_t st) {
    hipLaunchKernelGGL(acceleration_kernel,
                       dim3((unsigned)((g.nz + 63) / 64), (unsigned)((g.ny + 3) / 4), (unsigned)g.nx),
                       dim3(256), 0, st, g, phi_r, phi_b, gx, gy, gz);
    return hipGetLastError();
}

static dim3 residual_grid(const GridDims& g) {
    return dim3((unsigned)((g.nzh + 63) / 64), (unsigned)((g.ny + SOR_ROWS - 1) / SOR_ROWS),
                (unsigned)((g.nx + RES_PLANES - 1) / RES_PLANES));
}

long residual_scratch_len(const GridDims& g) {
    const dim3 d = residual_grid(g);
    return 2L * d.x * d.y * d.z;
}

hipError_t launch_residual(const GridDims& g, const SorCoef& c, const double* phi_r,
                           const double* phi_b, const double* rhs_r, const double* rhs_b,
                           float* r_r, float* r_b, double* part, double* acc, hipStream_t st) {
    const dim3 grid = residual_grid(g);
    hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, st, g, c, phi_r, phi_b, rhs_r, rhs_b,
                       r_r, r_b, part);
    hipLaunchKernelGGL(residual_finish_kernel, dim3(1), dim3(256), 0, st, part,
                       (long)grid.x * grid.y * grid.z, acc);
    return hipGetLastError();
}

hipError_t launch_start_residual_sq(const GridDims& g, const SorCoef& c, const float* e_r,
                                   const float* e_b, const float* r_r, const float* r_b,
                                   double* part, double* acc, hipStream_t st) {
    const dim3 grid = residual_grid(g);
    hipLaunchKernelGGL(start_residual_sq_kernel, grid, dim3(256), 0, st, g, (float)c.dx2i,
                       (float)c.dy2i, (float)c.dz2i, (float)c.factor, e_r, e_b, r_r, r_b, part);
    hipLaunchKernelGGL(residual_finish_kernel, dim3(1), dim3(256), 0, st, part,
                       (long)grid.x * grid.y * grid.z, acc);
    return hipGetLastError();
}

hipError_t launch_sor_sweep_f32(const GridDims& g, const SorCoef& c, int colour, float* e_r,
                                float* e_b, const float* r_r, const float* r_b, hipStream_t st) {
    float* mine = colour ? e_b : e_r;
    const float* other = colour ? e_r : e_b;
    const float* rhs = colour ? r_b : r_r;
    if (g.nzh % 4 == 0 && (long)(g.nx - 2) * (g.ny - 2) * (g.nzh / 4) < (1L << 31)) {
        const unsigned gpr = (unsigned)(g.nzh / 4);
        const unsigned total = (unsigned)((long)(g.nx - 2) * (g.ny - 2) * gpr);
        hipLaunchKernelGGL(sor_sweep_f32x4_kernel, dim3((total + 255) / 256), dim3(256), 0, st, g,
                           (float)c.dx2i, (float)c.dy2i, (float)c.dz2i, (float)(1.0 / c.factor),
                           (float)c.omega, colour, mine, other, rhs, gpr, total);
    } else {
        hipLaunchKernelGGL(sor_sweep_f32_kernel, sor_grid(g), dim3(256), 0, st, g,
                           (float)c.dx2i, (float)c.dy2i, (float)c.dz2i, (float)(1.0 / c.factor),
                           (float)c.omega, colour, mine, other, rhs);
    }
    return hipGetLastError();
}

hipError_t launch_correct(const GridDims& g, double* phi_r, double* phi_b, float* e_r,
                          float* e_b, hipStream_t st) {
    const long n = (long)g.nx * g.ny * g.nzh;
    hipLaunchKernelGGL(correct_kernel, dim3(grid_for(n)), dim3(256), 0, st, g, phi_r, phi_b, e_r,
                       e_b);
    return hipGetLastError();
}

}  // namespace hyd
