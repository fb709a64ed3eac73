// sor_fused.hip -- two red-black SOR iterations per pass over the potential.
//
// A block owns a tile of FJ rows (j) x FK cells (k) and a chunk of x planes.
// It streams planes through an LDS window and runs the four half-sweeps of
// two iterations as a pipeline along x: at step p it updates the red cells
// of plane p (iteration 1), the black cells of plane p-1 (iteration 1), the
// red cells of plane p-2 (iteration 2) and the black cells of plane p-3
// (iteration 2), each stage reading the values the previous stage left.
// Around the tile every stage also updates a ring of cells, 3, 2, 1 and 0
// wide, so the tile's own cells see exactly the neighbours a global sweep
// would give them; rings overlap between blocks, which is why the pass
// reads one copy of the potential and writes another.
//
// Each update is the expression of gravity_solver_3d.c:642-708 in the same
// order (FP64: gs = (lap - rhs) / factor; the FP32 correction sweeps
// multiply by the reciprocal), and red-black SOR does not depend on the
// order of cells within a colour, so the result equals two iterations of
// sweeps over the whole grid bit for bit.  Per iteration this moves about
// 12 B per cell of HBM traffic (FP64) instead of the 32 B of two separate
// colour sweeps.
#include "device_common.hpp"
#include "gravity_kernels.hpp"

#ifndef SOR2_FJ
#define SOR2_FJ 16
#endif
#ifndef SOR2_FK
#define SOR2_FK 32
#endif

namespace hyd {

namespace {
constexpr int FJ = SOR2_FJ, FK = SOR2_FK;  // tile: rows j x cells k (FK even)
constexpr int RJ = FJ + 8;               // LDS rows: the tile and a 4-row ring
constexpr int RM = FK / 2 + 5;           // LDS entries per row and colour: the
                                         // FK + 8 cells from the odd k0 - 4 span
                                         // FK / 2 + 5 split entries
constexpr int NP = 6;                    // planes held: p-4 .. p+1
constexpr int FT = 256;                  // threads per block
constexpr int NLD = (2 * RJ * RM + FT - 1) / FT;  // plane entries loaded per thread
}  // namespace

template <class T>
struct SorFusedArgs {
    GridDims g;
    const T *in_r, *in_b;  // state before the two iterations
    T *out_r, *out_b;      // state after them (interior cells written)
    const T *rhs_r, *rhs_b;
    T dx2i, dy2i, dz2i;
    T factor;              // FP64: the divisor 2(dx^-2+dy^-2+dz^-2); FP32: its reciprocal
    T omega;
    int cx;                // x planes per block
    int ntj, ntk, nxc;
    // max-update tracking: iteration index of the first of the two, the
    // slots of SorCtl, and the stop tolerance
    SorCtl* ctl;
    int it;
    double tol;
};

__device__ __forceinline__ double sor_quot(double x, double f) { return x / f; }
__device__ __forceinline__ float sor_quot(float x, float f) { return x * f; }

// LDS keeps the red-black split of the device arrays: plane slot, colour,
// row, entry m - m0 (cell k = 2m + ((p + j + colour) & 1)).  A stage then
// reads consecutive entries, free of bank conflicts.
template <class T>
using Window = T[NP][2][RJ][RM];

// Global loads run one pipeline step ahead of their use: a plane is fetched
// into registers during the step before it is stored to LDS, and each stage
// fetches the right-hand side of its next plane while it updates the
// current one, so no stage waits on memory latency.
// Every global load is issued unconditionally (indices clamped into the
// grid; values that belong to no cell are never used), so the compiler can
// count the loads in flight and wait only for the one it needs.
template <class T>
__device__ __forceinline__ void plane_fetch(const SorFusedArgs<T>& a, int p, int j0, int m0,
                                            T (&v)[NLD]) {
    const GridDims& g = a.g;
    p = min(max(p, 0), g.nx - 1);
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int t = min(threadIdx.x + u * FT, 2 * RJ * RM - 1);
        const int c = t / (RJ * RM), r = t - c * (RJ * RM);
        const int lj = r / RM, lm = r - lj * RM;
        const int j = min(max(j0 - 4 + lj, 0), g.ny - 1);
        const int m = min(max(m0 + lm, 0), g.nzh - 1);
        v[u] = (c ? a.in_b : a.in_r)[((long)p * g.ny + j) * g.nzh + m];
    }
}

template <class T>
__device__ __forceinline__ void plane_store(Window<T>& w, int p, const T (&v)[NLD]) {
    const int s = p % NP;
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int t = threadIdx.x + u * FT;
        if (t < 2 * RJ * RM) {
            const int c = t / (RJ * RM), r = t - c * (RJ * RM);
            const int lj = r / RM, lm = r - lj * RM;
            w[s][c][lj][lm] = v[u];  // entries outside the grid hold junk no update reads
        }
    }
}

// Stage S (0: red, iteration 1; 1: black, it. 1; 2: red, it. 2; 3: black,
// it. 2) covers the tile widened by a ring of 3 - S cells; thread t owns
// entries t, t + FT, ... of the stage's (row, cell-of-colour) list.
template <int S>
struct Stage {
    static constexpr int R = 3 - S;
    static constexpr int WJ = FJ + 2 * R, HK = (FK + 2 * R) / 2;
    static constexpr int N = WJ * HK;
    static constexpr int U = (N + FT - 1) / FT;
    static constexpr int colour = S & 1;
};

// cell (j, k) and split entry m of list entry t of stage S on plane q
template <int S>
__device__ __forceinline__ void stage_cell(int t, int q, int j0, int k0, int& j, int& k, int& m,
                                           int& par) {
    using G = Stage<S>;
    const int rj = t / G::HK, h = t - rj * G::HK;
    j = j0 - G::R + rj;
    par = (q + j + G::colour) & 1;                // parity of k for this colour
    m = ((k0 - G::R - par + 1) >> 1) + h;         // first m with 2m + par >= k0 - R
    k = 2 * m + par;
}

template <class T, int S>
__device__ __forceinline__ void rhs_fetch(const SorFusedArgs<T>& a, int q, int j0, int k0,
                                          T (&rr)[Stage<S>::U]) {
    using G = Stage<S>;
    const GridDims& g = a.g;
    const T* rhs = G::colour ? a.rhs_b : a.rhs_r;
    q = min(max(q, 0), g.nx - 1);
#pragma unroll
    for (int u = 0; u < G::U; ++u) {
        const int t = min(threadIdx.x + u * FT, G::N - 1);
        int j, k, m, par;
        stage_cell<S>(t, q, j0, k0, j, k, m, par);
        j = min(max(j, 0), g.ny - 1);
        m = min(max(m, 0), g.nzh - 1);
        rr[u] = rhs[((long)q * g.ny + j) * g.nzh + m];
    }
}

template <class T, int S, bool TRACK>
__device__ __forceinline__ void sor2_stage(const SorFusedArgs<T>& a, Window<T>& w, int q, int j0,
                                           int k0, int m0, int i0, int i1,
                                           const T (&rr)[Stage<S>::U], unsigned long long& rmax) {
    using G = Stage<S>;
    constexpr int c = G::colour, o = 1 - G::colour;
    const GridDims& g = a.g;
    const int sm = q % NP, sl = (q - 1) % NP, sh = (q + 1) % NP;
#pragma unroll
    for (int u = 0; u < G::U; ++u) {
        const int t = threadIdx.x + u * FT;
        int j, k, m, par;
        stage_cell<S>(t, q, j0, k0, j, k, m, par);
        if (t >= G::N || j < 1 || j > g.ny - 2 || k < 1 || k > g.nz - 2) continue;
        const int lj = j - j0 + 4, lm = m - m0;
        const T old = w[sm][c][lj][lm];
        // z neighbours k - 1, k + 1 are entries m - 1 + par and m + par
        const T lap = (w[sl][o][lj][lm] + w[sh][o][lj][lm]) * a.dx2i +
                      (w[sm][o][lj - 1][lm] + w[sm][o][lj + 1][lm]) * a.dy2i +
                      (w[sm][o][lj][lm - 1 + par] + w[sm][o][lj][lm + par]) * a.dz2i;
        const T gs = sor_quot(lap - rr[u], a.factor);
        w[sm][c][lj][lm] = old + a.omega * (gs - old);
        if (TRACK && j >= j0 && j < j0 + FJ && k >= k0 && k < k0 + FK && q >= i0 && q < i1) {
            const unsigned long long b = dbits(fabs((double)(gs - old)));
            rmax = b > rmax ? b : rmax;
        }
    }
}

template <class T, bool TRACK>
__global__ void __launch_bounds__(FT) sor2_kernel(SorFusedArgs<T> a) {
    __shared__ Window<T> w;
    const GridDims& g = a.g;
    if (TRACK) {
        // the stop rule of the single sweeps (sor_stop_check): nothing runs
        // once iteration it-1 or it-2 met the tolerance
        // (every wavefront reads the slots itself: all lanes are active here)
        bool stop = a.ctl->converged != 0;
        double prev = 0.0;
        if (!stop && a.it > 0) {
            prev = bitsd(sor_wave_slot_max(a.ctl, (a.it - 1) & 3));
            stop = prev < a.tol;
        }
        if (!stop && a.it > 1) {
            prev = bitsd(sor_wave_slot_max(a.ctl, (a.it - 2) & 3));
            stop = prev < a.tol;
        }
        if (stop) {
            if (blockIdx.x == 0 && threadIdx.x == 0 && !a.ctl->converged) {
                a.ctl->converged = 1;
                a.ctl->iters = a.it;
                a.ctl->last_res = prev;
            }
            return;
        }
    }
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);  // contiguous tiles per XCD
    const int tk = b % a.ntk, tj = (b / a.ntk) % a.ntj, xc = b / (a.ntk * a.ntj);
    const int j0 = 1 + tj * FJ, k0 = 1 + tk * FK;
    const int m0 = (k0 - 4) >> 1;  // first split entry held (k0 is odd)
    const int i0 = 1 + xc * a.cx, i1 = min(g.nx - 1, i0 + a.cx);
    unsigned long long r1 = 0ull, r2 = 0ull;

    // plane ranges: stage S updates planes [lo_S, hi_S]; loads cover [0, ld_hi]
    const int lo0 = max(1, i0 - 3), hi0 = min(g.nx - 2, i1 + 2);
    const int lo1 = max(1, i0 - 2), hi1 = min(g.nx - 2, i1 + 1);
    const int lo2 = max(1, i0 - 1), hi2 = min(g.nx - 2, i1);
    const int lo3 = i0, hi3 = i1 - 1;
    const int ld_hi = min(g.nx - 1, i1 + 3);
    const int pbeg = i0 - 3, pend = i1 + 2;

    T pv[NLD];
    T rr0[Stage<0>::U], rr1[Stage<1>::U], rr2[Stage<2>::U], rr3[Stage<3>::U];
    // prologue: planes pbeg-1 .. pbeg+1 in LDS, each stage's first right-hand side
    for (int p = max(0, pbeg - 1); p <= min(ld_hi, pbeg + 1); ++p) {
        plane_fetch(a, p, j0, m0, pv);
        plane_store(w, p, pv);
    }
    rhs_fetch<T, 0>(a, pbeg, j0, k0, rr0);
    rhs_fetch<T, 1>(a, pbeg - 1, j0, k0, rr1);
    rhs_fetch<T, 2>(a, pbeg - 2, j0, k0, rr2);
    rhs_fetch<T, 3>(a, pbeg - 3, j0, k0, rr3);

    for (int p = pbeg; p <= pend; ++p) {
        plane_fetch(a, p + 2, j0, m0, pv);
        __syncthreads();  // plane p + 1 stored, last step's write-out done
        if (p >= lo0 && p <= hi0) sor2_stage<T, 0, TRACK>(a, w, p, j0, k0, m0, i0, i1, rr0, r1);
        rhs_fetch<T, 0>(a, p + 1, j0, k0, rr0);
        __syncthreads();
        int q = p - 1;
        if (q >= lo1 && q <= hi1) sor2_stage<T, 1, TRACK>(a, w, q, j0, k0, m0, i0, i1, rr1, r1);
        rhs_fetch<T, 1>(a, q + 1, j0, k0, rr1);
        __syncthreads();
        q = p - 2;
        if (q >= lo2 && q <= hi2) sor2_stage<T, 2, TRACK>(a, w, q, j0, k0, m0, i0, i1, rr2, r2);
        rhs_fetch<T, 2>(a, q + 1, j0, k0, rr2);
        __syncthreads();
        q = p - 3;
        if (q >= lo3 && q <= hi3) sor2_stage<T, 3, TRACK>(a, w, q, j0, k0, m0, i0, i1, rr3, r2);
        rhs_fetch<T, 3>(a, q + 1, j0, k0, rr3);
        __syncthreads();
        if (q >= lo3 && q <= hi3) {
            // the tile's cells of plane q, both colours: FK / 2 entries per row
            const int s = q % NP;
            for (int t = threadIdx.x; t < 2 * FJ * (FK / 2); t += FT) {
                const int c = t / (FJ * (FK / 2)), r = t - c * (FJ * (FK / 2));
                const int rj = r / (FK / 2), h = r - rj * (FK / 2);
                const int j = j0 + rj;
                const int par = (q + j + c) & 1;
                const int m = ((k0 - par + 1) >> 1) + h;
                const int k = 2 * m + par;
                if (j > g.ny - 2 || k > g.nz - 2) continue;
                (c ? a.out_b : a.out_r)[((long)q * g.ny + j) * g.nzh + m] =
                    w[s][c][rj + 4][m - m0];
            }
        }
        // plane p + 2 takes the slot of plane p - 4, which B2 of this step
        // was the last to read (before the barrier above)
        if (p + 2 >= 0 && p + 2 <= ld_hi) plane_store(w, p + 2, pv);
    }
    if (TRACK) {
        r1 = block_max_u64(r1);
        __syncthreads();
        r2 = block_max_u64(r2);
        if (threadIdx.x == 0) {
            sor_slot_add(a.ctl, a.it & 3, blockIdx.x, r1);
            sor_slot_add(a.ctl, (a.it + 1) & 3, blockIdx.x, r2);
        }
    }
}

// ----------------------------------------------------------- launchers --
static int sor2_cx(const GridDims& g, long tiles, int blocks_per_cu) {
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
    const int nin = g.nx - 2;
    const long slots = (long)ncu * blocks_per_cu;
    int best_cx = nin;
    long best = -1;
    for (int cx = 8; cx <= 128; ++cx) {
        const long nxc = (nin + cx - 1) / cx;
        const long rounds = (tiles * nxc + slots - 1) / slots;
        const long cost = rounds * (cx + 6);  // 6 pipeline fill/drain steps
        if (best < 0 || cost < best) {
            best = cost;
            best_cx = cx;
        }
    }
    return best_cx < nin ? best_cx : nin;
}

template <class T, bool TRACK>
static hipError_t launch_sor2_t(SorFusedArgs<T> a, hipStream_t st) {
    const GridDims& g = a.g;
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return hipErrorInvalidValue;
    a.ntj = (g.ny - 2 + FJ - 1) / FJ;
    a.ntk = (g.nz - 2 + FK - 1) / FK;
    static int bpc = 0;
    if (bpc == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, sor2_kernel<T, TRACK>, FT, 0) !=
                hipSuccess ||
            v < 1)
            v = 1;
        bpc = v;
    }
    a.cx = sor2_cx(g, (long)a.ntj * a.ntk, bpc);
    a.nxc = (g.nx - 2 + a.cx - 1) / a.cx;
    const long nblk = (long)a.ntj * a.ntk * a.nxc;
    hipLaunchKernelGGL((sor2_kernel<T, TRACK>), dim3((unsigned)nblk), dim3(FT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sor2(const GridDims& g, const SorCoef& c, const double* in_r,
                       const double* in_b, double* out_r, double* out_b, const double* rhs_r,
                       const double* rhs_b, SorCtl* ctl, int it, double tol, hipStream_t st) {
    SorFusedArgs<double> a;
    a.g = g;
    a.in_r = in_r; a.in_b = in_b; a.out_r = out_r; a.out_b = out_b;
    a.rhs_r = rhs_r; a.rhs_b = rhs_b;
    a.dx2i = c.dx2i; a.dy2i = c.dy2i; a.dz2i = c.dz2i;
    a.factor = c.factor;
    a.omega = c.omega;
    a.ctl = ctl;
    a.it = it;
    a.tol = tol;
    if (ctl) {
        const hipError_t e = launch_sor_prepare(ctl, it, st);
        if (e != hipSuccess) return e;
        return launch_sor2_t<double, true>(a, st);
    }
    return launch_sor2_t<double, false>(a, st);
}

hipError_t launch_sor2_f32(const GridDims& g, const SorCoef& c, const float* in_r,
                           const float* in_b, float* out_r, float* out_b, const float* rhs_r,
                           const float* rhs_b, hipStream_t st) {
    SorFusedArgs<float> a;
    a.g = g;
    a.in_r = in_r; a.in_b = in_b; a.out_r = out_r; a.out_b = out_b;
    a.rhs_r = rhs_r; a.rhs_b = rhs_b;
    a.dx2i = (float)c.dx2i; a.dy2i = (float)c.dy2i; a.dz2i = (float)c.dz2i;
    a.factor = (float)(1.0 / c.factor);
    a.omega = (float)c.omega;
    a.ctl = nullptr;
    a.it = 0;
    a.tol = 0.0;
    return launch_sor2_t<float, false>(a, st);
}

}  // namespace hyd
