// multigrid.hip -- FP32 multigrid V-cycle kernels for the correction
// equation of the defect-correction Poisson solve (see multigrid.hpp for the
// grid hierarchy).  Every kernel runs on a 3D launch grid: 64 lanes along
// the fastest index, 4 rows (j) per block, one plane (i) per blockIdx.z, so
// no index needs an integer division.  Sums are taken in a fixed order, so
// a V-cycle is reproducible run to run.
#include "multigrid.hpp"

namespace hyd {

namespace {

constexpr int MG_ROWS = 4;

__device__ __forceinline__ long split_at(const GridDims& g, int i, int j, int k) {
    return ((long)i * g.ny + j) * g.nzh + (k >> 1);
}

// Interpolation along one direction: position xs on the finer lattice (the
// coarser level's point c at xs = 2c).  Between two coarser points the weights
// are 1/2 each, except at xs = 1, whose left point is the boundary (value 0)
// and whose weight from point 1 is pw1.
__device__ __forceinline__ void interp_1d(int xs, float pw1, int& c0, int& c1, float& w0,
                                          float& w1) {
    c0 = xs >> 1;
    if (xs & 1) {
        c1 = c0 + 1;
        w0 = (xs == 1) ? 0.0f : 0.5f;
        w1 = (xs == 1) ? pw1 : 0.5f;
    } else {
        c1 = c0;
        w0 = 1.0f;
        w1 = 0.0f;
    }
}

__device__ __forceinline__ float interpolate(const float* __restrict__ u, int ny, int nz, int xs,
                                             int ys, int zs, float px, float py, float pz) {
    int x0, x1, y0, y1, z0, z1;
    float wx0, wx1, wy0, wy1, wz0, wz1;
    interp_1d(xs, px, x0, x1, wx0, wx1);
    interp_1d(ys, py, y0, y1, wy0, wy1);
    interp_1d(zs, pz, z0, z1, wz0, wz1);
    const long a0 = ((long)x0 * ny + y0) * nz, a1 = ((long)x0 * ny + y1) * nz;
    const long b0 = ((long)x1 * ny + y0) * nz, b1 = ((long)x1 * ny + y1) * nz;
    const float v0 = wy0 * (wz0 * u[a0 + z0] + wz1 * u[a0 + z1]) +
                     wy1 * (wz0 * u[a1 + z0] + wz1 * u[a1 + z1]);
    const float v1 = wy0 * (wz0 * u[b0 + z0] + wz1 * u[b0 + z1]) +
                     wy1 * (wz0 * u[b1 + z0] + wz1 * u[b1 + z1]);
    return wx0 * v0 + wx1 * v1;
}

// Restriction weight along one direction of coarse index c from the finer
// point at offset d (-1, 0, 1).
__device__ __forceinline__ float rweight(const MgAxis& A, int c, int d) {
    if (c == 1) return d < 0 ? A.r1[0] : (d == 0 ? A.r1[1] : A.r1[2]);
    return d ? 0.25f : 0.5f;
}

// The operator along one direction at index i: off-diagonal part.
__device__ __forceinline__ float axis_off(const MgAxis& A, int i, float um, float up) {
    return (i == 1) ? A.am1 * um + A.ap1 * up : A.a * (um + up);
}

__device__ __forceinline__ float axis_diag(const MgAxis& A, int i) {
    return (i == 1) ? A.am1 + A.ap1 : 2.0f * A.a;
}

// ------------------------------------------------------------- level 0 --
// q = r - L e on the split grid (both colours of entry m); 0 on the boundary.
__global__ void __launch_bounds__(256) mg_fine_residual_kernel(
    GridDims g, float dx2i, float dy2i, float dz2i, float factor, const float* __restrict__ er,
    const float* __restrict__ eb, const float* __restrict__ rr, const float* __restrict__ rb,
    float* __restrict__ qr, float* __restrict__ qb) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (m >= g.nzh || j >= g.ny) return;
    const long sxh = (long)g.ny * g.nzh;
    const long base = ((long)i * g.ny + j) * g.nzh;
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
        const int k = 2 * m + ((i + j + colour) & 1);
        if (k >= g.nz) continue;
        float* q = colour ? qb : qr;
        if (i < 1 || i > g.nx - 2 || j < 1 || j > g.ny - 2 || k < 1 || k > g.nz - 2) {
            q[base + m] = 0.0f;
            continue;
        }
        const float* mine = colour ? eb : er;
        const float* other = colour ? er : eb;
        const float lap = (other[base - sxh + m] + other[base + sxh + m]) * dx2i +
                          (other[base - g.nzh + m] + other[base + g.nzh + m]) * dy2i +
                          (other[base + ((k - 1) >> 1)] + other[base + ((k + 1) >> 1)]) * dz2i -
                          factor * mine[base + m];
        q[base + m] = (colour ? rb : rr)[base + m] - lap;
    }
}

// Restriction onto level 1: coarse point c gathers fine 2c-2 .. 2c per
// direction (weights 1/4, 1/2, 1/4; r1 at c = 1).  Also clears u1.
__global__ void __launch_bounds__(256) mg_restrict_fine_kernel(GridDims g,
                                                               const float* __restrict__ qr,
                                                               const float* __restrict__ qb,
                                                               MgLevel c, float* __restrict__ f1,
                                                               float* __restrict__ u1) {
    const int K = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int I = blockIdx.z;
    if (K >= c.nz || J >= c.ny) return;
    const long t = ((long)I * c.ny + J) * c.nz + K;
    float f = 0.0f;
    if (I >= 1 && I <= c.nx - 2 && J >= 1 && J <= c.ny - 2 && K >= 1 && K <= c.nz - 2) {
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = 2 * I - 1 + dx;
            const float wx = rweight(c.ax[0], I, dx);
            for (int dy = -1; dy <= 1; ++dy) {
                const int y = 2 * J - 1 + dy;
                const float wxy = wx * rweight(c.ax[1], J, dy);
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const int z = 2 * K - 1 + dz;
                    const float* q = ((x + y + z) & 1) ? qb : qr;
                    f += (wxy * rweight(c.ax[2], K, dz)) * q[split_at(g, x, y, z)];
                }
            }
        }
    }
    f1[t] = f;
    u1[t] = 0.0f;
}

// e += P u1 on the interior of the split grid (fine x sits at 2c - 1).
__global__ void __launch_bounds__(256) mg_prolong_fine_kernel(GridDims g, MgLevel c,
                                                              const float* __restrict__ u1,
                                                              float* __restrict__ er,
                                                              float* __restrict__ eb) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (m >= g.nzh || j < 1 || j > g.ny - 2 || i < 1 || i > g.nx - 2) return;
    const long base = ((long)i * g.ny + j) * g.nzh;
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
        const int k = 2 * m + ((i + j + colour) & 1);
        if (k < 1 || k > g.nz - 2) continue;
        float* e = colour ? eb : er;
        e[base + m] += interpolate(u1, c.ny, c.nz, i + 1, j + 1, k + 1, 0.5f, 0.5f, 0.5f);
    }
}

// ------------------------------------------------------- coarse levels --
// Point operations shared by the one-phase-per-launch kernels and the
// one-workgroup tail kernel, so both give the same bits.  T is the index
// type (long for global arrays, int inside the tail's LDS).
template <class T>
__device__ __forceinline__ void smooth_at(const MgLevel& L, int i, int j, int k, float omega,
                                          float* u, const float* f) {
    const T sx = (T)L.ny * L.nz;
    const T t = ((T)i * L.ny + j) * L.nz + k;
    const float off = axis_off(L.ax[0], i, u[t - sx], u[t + sx]) +
                      axis_off(L.ax[1], j, u[t - L.nz], u[t + L.nz]) +
                      axis_off(L.ax[2], k, u[t - 1], u[t + 1]);
    const float diag = axis_diag(L.ax[0], i) + axis_diag(L.ax[1], j) + axis_diag(L.ax[2], k);
    const float old = u[t];
    const float gs = (off - f[t]) / diag;
    u[t] = old + omega * (gs - old);
}

// f - L u at an interior point
template <class T>
__device__ __forceinline__ float residual_at(const MgLevel& L, int i, int j, int k,
                                             const float* u, const float* f) {
    const T sx = (T)L.ny * L.nz;
    const T t = ((T)i * L.ny + j) * L.nz + k;
    const float off = axis_off(L.ax[0], i, u[t - sx], u[t + sx]) +
                      axis_off(L.ax[1], j, u[t - L.nz], u[t + L.nz]) +
                      axis_off(L.ax[2], k, u[t - 1], u[t + 1]);
    const float diag = axis_diag(L.ax[0], i) + axis_diag(L.ax[1], j) + axis_diag(L.ax[2], k);
    return f[t] - (off - diag * u[t]);
}

// restriction of q (level F) into interior point (I, J, K) of level C
template <class T>
__device__ __forceinline__ float restrict_at(const MgLevel& F, const float* q, const MgLevel& C,
                                             int I, int J, int K) {
    float f = 0.0f;
    for (int dx = -1; dx <= 1; ++dx) {
        const float wx = rweight(C.ax[0], I, dx);
        for (int dy = -1; dy <= 1; ++dy) {
            const float wxy = wx * rweight(C.ax[1], J, dy);
            const T row = ((T)(2 * I + dx) * F.ny + (2 * J + dy)) * F.nz + 2 * K;
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) f += (wxy * rweight(C.ax[2], K, dz)) * q[row + dz];
        }
    }
    return f;
}

__global__ void __launch_bounds__(256) mg_smooth_kernel(MgLevel L, int colour, float omega,
                                                        float* __restrict__ u,
                                                        const float* __restrict__ f) {
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int i = 1 + blockIdx.z;
    const int k = 2 * m + ((i + j + colour) & 1);
    if (j > L.ny - 2 || k < 1 || k > L.nz - 2) return;
    smooth_at<long>(L, i, j, k, omega, u, f);
}

__global__ void __launch_bounds__(256) mg_residual_kernel(MgLevel L, const float* __restrict__ u,
                                                          const float* __restrict__ f,
                                                          float* __restrict__ q) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= L.nz || j >= L.ny) return;
    const long t = ((long)i * L.ny + j) * L.nz + k;
    const bool inner = i >= 1 && i <= L.nx - 2 && j >= 1 && j <= L.ny - 2 && k >= 1 &&
                       k <= L.nz - 2;
    q[t] = inner ? residual_at<long>(L, i, j, k, u, f) : 0.0f;
}

// Restriction between coarse levels: coarse c gathers fine 2c-1 .. 2c+1.
__global__ void __launch_bounds__(256) mg_restrict_kernel(MgLevel F, const float* __restrict__ q,
                                                          MgLevel C, float* __restrict__ fc,
                                                          float* __restrict__ uc) {
    const int K = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int I = blockIdx.z;
    if (K >= C.nz || J >= C.ny) return;
    const long t = ((long)I * C.ny + J) * C.nz + K;
    const bool inner = I >= 1 && I <= C.nx - 2 && J >= 1 && J <= C.ny - 2 && K >= 1 &&
                       K <= C.nz - 2;
    fc[t] = inner ? restrict_at<long>(F, q, C, I, J, K) : 0.0f;
    uc[t] = 0.0f;
}

__global__ void __launch_bounds__(256) mg_prolong_kernel(MgLevel C, const float* __restrict__ uc,
                                                         MgLevel F, float* __restrict__ uf) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * MG_ROWS + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k < 1 || k > F.nz - 2 || j < 1 || j > F.ny - 2 || i < 1 || i > F.nx - 2) return;
    const long t = ((long)i * F.ny + j) * F.nz + k;
    uf[t] += interpolate(uc, C.ny, C.nz, i, j, k, F.ax[0].pw1, F.ax[1].pw1, F.ax[2].pw1);
}

// ---------------------------------------------------------------- tail --
// The V-cycle from level `first` of the tail table down to its coarsest
// level in one workgroup, with every level's u, f and q in LDS: the same
// phases and point operations as the launch-per-phase path, separated by
// barriers instead of launches.  Reads f of the first level from global
// memory (u starts at 0) and writes its u back.
constexpr int MG_TAIL_THREADS = 1024;

__device__ void tail_smooth(const MgLevel& L, int colour, float omega, float* u, const float* f) {
    const int mh = (L.nz + 1) / 2, rows = L.ny - 2;
    const int total = (L.nx - 2) * rows * mh;
    for (int t = threadIdx.x; t < total; t += MG_TAIL_THREADS) {
        const int m = t % mh, r = t / mh;
        const int j = 1 + r % rows, i = 1 + r / rows;
        const int k = 2 * m + ((i + j + colour) & 1);
        if (k >= 1 && k <= L.nz - 2) smooth_at<int>(L, i, j, k, omega, u, f);
    }
    __syncthreads();
}

__device__ void tail_sweeps(const MgLevel& L, int n, float omega, float* u, const float* f) {
    for (int s = 0; s < n; ++s) {
        tail_smooth(L, 0, omega, u, f);
        tail_smooth(L, 1, omega, u, f);
    }
}

__global__ void __launch_bounds__(MG_TAIL_THREADS) mg_tail_kernel(const MgTail* __restrict__ tab,
                                                                  const float* __restrict__ f_in,
                                                                  float* __restrict__ u_out) {
    __shared__ float sm[MG_TAIL_LDS_MAX / sizeof(float)];
    const int nl = tab->count;
    {
        const MgLevel& L = tab->lv[0];
        const int n = L.nx * L.ny * L.nz;
        float* u = sm + tab->off_u[0];
        float* f = sm + tab->off_f[0];
        for (int t = threadIdx.x; t < n; t += MG_TAIL_THREADS) {
            f[t] = f_in[t];
            u[t] = 0.0f;
        }
        __syncthreads();
    }
    for (int l = 0; l + 1 < nl; ++l) {  // down
        const MgLevel& L = tab->lv[l];
        const MgLevel& C = tab->lv[l + 1];
        float* u = sm + tab->off_u[l];
        float* f = sm + tab->off_f[l];
        float* q = sm + tab->off_q[l];
        tail_sweeps(L, tab->nu1, tab->omega, u, f);
        const int n = L.nx * L.ny * L.nz;
        for (int t = threadIdx.x; t < n; t += MG_TAIL_THREADS) {
            const int k = t % L.nz, r = t / L.nz;
            const int j = r % L.ny, i = r / L.ny;
            const bool inner = i >= 1 && i <= L.nx - 2 && j >= 1 && j <= L.ny - 2 && k >= 1 &&
                               k <= L.nz - 2;
            q[t] = inner ? residual_at<int>(L, i, j, k, u, f) : 0.0f;
        }
        __syncthreads();
        float* uc = sm + tab->off_u[l + 1];
        float* fc = sm + tab->off_f[l + 1];
        const int nc = C.nx * C.ny * C.nz;
        for (int t = threadIdx.x; t < nc; t += MG_TAIL_THREADS) {
            const int K = t % C.nz, r = t / C.nz;
            const int J = r % C.ny, I = r / C.ny;
            const bool inner = I >= 1 && I <= C.nx - 2 && J >= 1 && J <= C.ny - 2 && K >= 1 &&
                               K <= C.nz - 2;
            fc[t] = inner ? restrict_at<int>(L, q, C, I, J, K) : 0.0f;
            uc[t] = 0.0f;
        }
        __syncthreads();
    }
    {
        const MgLevel& L = tab->lv[nl - 1];  // coarsest: Gauss-Seidel sweeps
        tail_sweeps(L, max(L.nx, max(L.ny, L.nz)), 1.0f, sm + tab->off_u[nl - 1],
                    sm + tab->off_f[nl - 1]);
    }
    for (int l = nl - 2; l >= 0; --l) {  // up
        const MgLevel& L = tab->lv[l];
        const MgLevel& C = tab->lv[l + 1];
        float* u = sm + tab->off_u[l];
        const float* uc = sm + tab->off_u[l + 1];
        const int mz = L.nz - 2, my = L.ny - 2;
        const int n = (L.nx - 2) * my * mz;
        for (int t = threadIdx.x; t < n; t += MG_TAIL_THREADS) {
            const int k = 1 + t % mz, r = t / mz;
            const int j = 1 + r % my, i = 1 + r / my;
            u[(i * L.ny + j) * L.nz + k] +=
                interpolate(uc, C.ny, C.nz, i, j, k, L.ax[0].pw1, L.ax[1].pw1, L.ax[2].pw1);
        }
        __syncthreads();
        tail_sweeps(L, tab->nu2, tab->omega, u, sm + tab->off_f[l]);
    }
    {
        const MgLevel& L = tab->lv[0];
        const int n = L.nx * L.ny * L.nz;
        const float* u = sm + tab->off_u[0];
        for (int t = threadIdx.x; t < n; t += MG_TAIL_THREADS) u_out[t] = u[t];
    }
}

dim3 grid3(int fast, int rows, int planes) {
    return dim3((unsigned)((fast + 63) / 64), (unsigned)((rows + MG_ROWS - 1) / MG_ROWS),
                (unsigned)planes);
}

}  // namespace

// ----------------------------------------------------------- launchers --
hipError_t launch_mg_fine_residual(const GridDims& g, const SorCoef& c, const float* e_r,
                                   const float* e_b, const float* r_r, const float* r_b,
                                   float* q_r, float* q_b, hipStream_t st) {
    hipLaunchKernelGGL(mg_fine_residual_kernel, grid3(g.nzh, g.ny, g.nx), dim3(256), 0, st, g,
                       (float)c.dx2i, (float)c.dy2i, (float)c.dz2i, (float)c.factor, e_r, e_b,
                       r_r, r_b, q_r, q_b);
    return hipGetLastError();
}

hipError_t launch_mg_restrict_fine(const GridDims& g, const float* q_r, const float* q_b,
                                   const MgLevel& c1, float* f1, float* u1, hipStream_t st) {
    if (c1.nx != g.nx / 2 + 1 || c1.ny != g.ny / 2 + 1 || c1.nz != g.nz / 2 + 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mg_restrict_fine_kernel, grid3(c1.nz, c1.ny, c1.nx), dim3(256), 0, st, g,
                       q_r, q_b, c1, f1, u1);
    return hipGetLastError();
}

hipError_t launch_mg_prolong_fine(const GridDims& g, const MgLevel& c1, const float* u1,
                                  float* e_r, float* e_b, hipStream_t st) {
    if (c1.nx != g.nx / 2 + 1 || c1.ny != g.ny / 2 + 1 || c1.nz != g.nz / 2 + 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mg_prolong_fine_kernel, grid3(g.nzh, g.ny, g.nx), dim3(256), 0, st, g, c1,
                       u1, e_r, e_b);
    return hipGetLastError();
}

hipError_t launch_mg_smooth(const MgLevel& L, int colour, float omega, float* u, const float* f,
                            hipStream_t st) {
    hipLaunchKernelGGL(mg_smooth_kernel, grid3((L.nz + 1) / 2, L.ny - 2, L.nx - 2), dim3(256), 0,
                       st, L, colour, omega, u, f);
    return hipGetLastError();
}

hipError_t launch_mg_residual(const MgLevel& L, const float* u, const float* f, float* q,
                              hipStream_t st) {
    hipLaunchKernelGGL(mg_residual_kernel, grid3(L.nz, L.ny, L.nx), dim3(256), 0, st, L, u, f, q);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(const MgLevel& fine, const float* q, const MgLevel& coarse,
                              float* fc, float* uc, hipStream_t st) {
    if (fine.nx != 2 * coarse.nx - 1 || fine.ny != 2 * coarse.ny - 1 ||
        fine.nz != 2 * coarse.nz - 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mg_restrict_kernel, grid3(coarse.nz, coarse.ny, coarse.nx), dim3(256), 0,
                       st, fine, q, coarse, fc, uc);
    return hipGetLastError();
}

hipError_t launch_mg_tail(const MgTail* d_tab, int lds_floats, const float* f_first,
                          float* u_first, hipStream_t st) {
    if (lds_floats < 1 || sizeof(float) * (size_t)lds_floats > MG_TAIL_LDS_MAX)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mg_tail_kernel, dim3(1), dim3(MG_TAIL_THREADS), 0, st, d_tab, f_first,
                       u_first);
    return hipGetLastError();
}

hipError_t launch_mg_prolong(const MgLevel& coarse, const float* uc, const MgLevel& fine,
                             float* uf, hipStream_t st) {
    if (fine.nx != 2 * coarse.nx - 1 || fine.ny != 2 * coarse.ny - 1 ||
        fine.nz != 2 * coarse.nz - 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mg_prolong_kernel, grid3(fine.nz, fine.ny, fine.nx), dim3(256), 0, st,
                       coarse, uc, fine, uf);
    return hipGetLastError();
}

}  // namespace hyd
