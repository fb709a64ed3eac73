// nn_kernels.hip -- kernels of the learned initial guess for the Poisson
// solve: BF16 3x3x3 convolutions on MFMA (forward, input gradient through
// the same kernel with the adjoint weights, weight gradient), pooling, the
// loss and Adam.  No reference counterpart: the reference starts SOR from the
// previous potential or from zero (gravity_solver_3d.c:87, :566).
//
// MFMA operand layout of v_mfma_f32_16x16x16bf16_1k (wave64): lane l holds
// A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16] for i = 0..3 and
// receives D[4 (l / 16) + i][l % 16]; the forward and backward tests against
// a CPU reference fail on any other layout.
#include "nn_kernels.hpp"

namespace hyd {
namespace nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

// FP32 -> BF16 rounded to nearest even: gfx950's v_cvt_pk_bf16_f32, two
// values per instruction (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v2f{lo, hi}, v2bf));
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

__device__ __forceinline__ void tap_offset(int t, int& dx, int& dy, int& dz) {
    dx = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dz = t % 3 - 1;
}

// ------------------------------------------------------------- forward --
// The convolution with the input tile staged in LDS.  A block owns a
// 4 (x) x 4 (y) x 16 (z) brick of output voxels at a time; wavefront w takes
// x slice w and its four y rows as four 16-voxel tiles along z.  The product
// is formed as D = W^T X^T (output channels down, voxels across), so that a
// lane ends with four consecutive channels of one voxel: one 8-byte store.  The
// brick's input, widened by one voxel on every side (upsampled on the fly
// for the decoder layers, zero outside the grid), is read from HBM once into
// LDS, so every A fragment comes from LDS instead of a 27-fold global gather.
// The layer's packed B fragments are staged in LDS once per block, and the
// block walks bricks blockIdx.x, blockIdx.x + gridDim.x, ...
constexpr int BX = 4, BY = 4, BZ = 16;
constexpr int LX = BX + 2, LY = BY + 2, LZ = BZ + 2;
#ifndef HYD_CONV_MAX_BLOCKS
#define HYD_CONV_MAX_BLOCKS 2048  // 8 per CU; 1024 and 4096 measured 2-4% slower at 128^3
#endif
constexpr int CONV_MAX_BLOCKS = HYD_CONV_MAX_BLOCKS;

template <int CINP, int NTL>
__global__ void __launch_bounds__(256, (CINP == 12 ? 3 : 1))  // 12: 3 waves per SIMD, no spills
conv_fwd_lds_kernel(ConvArgs a) {
    constexpr int KSTEPS = (27 * CINP + 15) / 16;
    constexpr int G4 = CINP / 4;  // 4-channel groups per voxel
    __shared__ v4s s_in[LX * LY * LZ * G4];
    __shared__ v4s s_w[NTL * KSTEPS * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = a.n;
    const int nbz = (n + BZ - 1) / BZ, nby = (n + BY - 1) / BY, nbx = (n + BX - 1) / BX;
    const int nbricks = nbx * nby * nbz;
    const int nin = a.up ? n / 2 : n;
    const v4s* xin = reinterpret_cast<const v4s*>(a.x);
    const v4s* wp = reinterpret_cast<const v4s*>(a.wp);
    for (int e = threadIdx.x; e < NTL * KSTEPS * 64; e += 256) s_w[e] = wp[e];
    const int g = lane >> 4, m = lane & 15;

    // The input brick, (x, y, z, group) entries with z and group fastest, is
    // read into registers one brick ahead: the loads of brick b + gridDim.x
    // are in flight while the MFMAs of brick b run.
    constexpr int NE = LX * LY * LZ * G4;
    constexpr int PER = (NE + 255) / 256;
    v4s stage[PER];
    auto load_brick = [&](int blk) {
        const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            int e = threadIdx.x + 256 * it;
            // From 12 channels up the indices are recomputed per brick: kept
            // live across the brick loop they cost a wave per SIMD (12
            // channels at 128^3: 119 us at 2 waves, 102 us at 3).  With 4
            // and 8 channels occupancy does not change and recomputing only
            // adds VALU work.
            if constexpr (CINP >= 12) asm volatile("" : "+v"(e));
            const int q = e % G4, vz = (e / G4) % LZ, vy = (e / (G4 * LZ)) % LY,
                      vx = e / (G4 * LZ * LY);
            const int ii = bx * BX - 1 + vx, jj = by * BY - 1 + vy, kz = bz * BZ - 1 + vz;
            stage[it] = v4s{0, 0, 0, 0};
            if (e < NE && (unsigned)ii < (unsigned)n && (unsigned)jj < (unsigned)n &&
                (unsigned)kz < (unsigned)n) {
                const int si = a.up ? ii >> 1 : ii, sj = a.up ? jj >> 1 : jj,
                          sk = a.up ? kz >> 1 : kz;
                stage[it] = xin[(((long)si * nin + sj) * nin + sk) * G4 + q];
            }
        }
    };
    if (blockIdx.x < nbricks) load_brick(blockIdx.x);

    for (int blk = blockIdx.x; blk < nbricks; blk += gridDim.x) {
        const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * BZ;
        __syncthreads();  // the previous brick's reads are done
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const int e = threadIdx.x + 256 * it;
            if (e < NE) s_in[e] = stage[it];
        }
        __syncthreads();
        if (blk + (int)gridDim.x < nbricks) load_brick(blk + gridDim.x);

        v4f acc[4][NTL];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) acc[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll  // taps, channels and LDS offsets become constants
        for (int ks = 0; ks < KSTEPS; ++ks) {
            const int kk = (ks * 4 + g) * 4;
            const int t = kk / CINP, ci = kk % CINP;
            int dx = 0, dy = 0, dz = 0;
            if (t < 27) tap_offset(t, dx, dy, dz);
            v4s bw[NTL];
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) bw[nt] = s_w[(nt * KSTEPS + ks) * 64 + lane];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                // voxel (x0 + wave, y0 + mt, z0 + m) shifted by the tap, in the brick
                v4s av = v4s{0, 0, 0, 0};
                if (t < 27)
                    av = s_in[(((wave + 1 + dx) * LY + (mt + 1 + dy)) * LZ + (m + 1 + dz)) * G4 +
                              ci / 4];
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt)
                    acc[mt][nt] =
                        __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw[nt], av, acc[mt][nt], 0, 0, 0);
            }
        }
        // epilogue: bias, ReLU, optional skip add.  D = W^T X^T: lane (g, m)
        // holds channels nt*16 + 4g .. +3 of voxel z0 + m, stored as 8 bytes.
        const int xi = x0 + wave, zk = z0 + m;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int yj = y0 + mt;
            if (xi >= n || yj >= n || zk >= n) continue;
            const long v = ((long)xi * n + yj) * n + zk;
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) {
                const int c0 = nt * 16 + 4 * g;
                if (c0 >= a.coutp) continue;
                float val[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float t = acc[mt][nt][i];
                    if (c0 + i < a.cout) {
                        if (a.bias) t += a.bias[c0 + i];
                        if (a.relu) t = t > 0.f ? t : 0.f;
                    } else {
                        t = 0.f;
                    }
                    val[i] = t;
                }
                if (a.mask) {  // zeroing commutes with the BF16 rounding below
                    const v2u mk = *reinterpret_cast<const v2u*>(a.mask + v * a.coutp + c0);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (!(bf_lo(mk[h]) > 0.f)) val[2 * h] = 0.f;
                        if (!(bf_hi(mk[h]) > 0.f)) val[2 * h + 1] = 0.f;
                    }
                }
                const v2u rb = v2u{f2bf2(val[0], val[1]), f2bf2(val[2], val[3])};
                if (a.yr) *reinterpret_cast<v2u*>(a.yr + v * a.coutp + c0) = rb;
                if (a.yd) {
                    v2u sk = v2u{0u, 0u};
                    if (a.skip) sk = *reinterpret_cast<const v2u*>(a.skip + v * a.coutp + c0);
                    v2u o;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        o[h] = f2bf2(bf_lo(rb[h]) + bf_lo(sk[h]), bf_hi(rb[h]) + bf_hi(sk[h]));
                    *reinterpret_cast<v2u*>(a.yd + v * a.coutp + c0) = o;
                }
                if (a.yf && c0 == 0) a.yf[v] = a.add ? val[0] + a.add[v] : val[0];
            }
        }
    }
}

template <int CINP>
void launch_fwd_cin(const ConvArgs& a, unsigned nb, hipStream_t st) {
    if (a.cout > 16)
        hipLaunchKernelGGL((conv_fwd_lds_kernel<CINP, 2>), dim3(nb), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((conv_fwd_lds_kernel<CINP, 1>), dim3(nb), dim3(256), 0, st, a);
}

hipError_t conv_forward(const ConvArgs& a, hipStream_t st) {
    if (a.cout < 1 || a.cout > 32 || a.coutp < a.cout || a.coutp % 4 || a.n < 1 ||
        (a.up && a.n % 2))
        return hipErrorInvalidValue;  // one wavefront covers at most two 16-channel tiles
    const long nbricks = (long)((a.n + BX - 1) / BX) * ((a.n + BY - 1) / BY) * ((a.n + BZ - 1) / BZ);
    const unsigned nb = (unsigned)(nbricks < CONV_MAX_BLOCKS ? nbricks : CONV_MAX_BLOCKS);
    switch (a.cinp) {
        case 4: launch_fwd_cin<4>(a, nb, st); break;
        case 8: launch_fwd_cin<8>(a, nb, st); break;
        case 12: launch_fwd_cin<12>(a, nb, st); break;
        case 16: launch_fwd_cin<16>(a, nb, st); break;
        case 24: launch_fwd_cin<24>(a, nb, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// -------------------------------------------------------- weight gradient --
// dW[t][ci][co] = sum_v x[v + off_t][ci] g[v][co] as D[co][(t, ci)] =
// G^T X on MFMA: M = output channels, N = rows (t, ci), K = voxels.  A block
// walks bricks of 4 x 4 x 16 voxels (the forward kernel's bricks).  Per brick
// it stages the gradient transposed, gs[co][voxel], and the input with a
// one-voxel halo as three copies shifted by dz = -1, 0, 1 along z, so that
// every fragment is one aligned 8-byte LDS read.  Both LDS tiles pad their
// rows by 16 elements (8 banks): unpadded, the 16 rows of a fragment fell on
// the same 8 of the 64 banks.  Rows past 27 * cinp are
// zero except the first, which is all ones and yields the bias gradient in
// the same product.  Each block writes its sums to part[block]; a second
// kernel adds the blocks in a fixed order (no atomics: the gradient does
// not depend on scheduling).
template <int CINP, int MT>
__global__ void __launch_bounds__(256) conv_wgrad_lds_kernel(WgradArgs a) {
    constexpr int NROWS = 27 * CINP;
    constexpr int NT = (NROWS + 1 + 15) / 16;  // row tiles, the ones row included
    constexpr int PAIRS = NT * MT;
    constexpr int PW = (PAIRS + 3) / 4;  // tile pairs per wavefront
    constexpr int G4 = CINP / 4;
    constexpr int XCH = LX * LY * BZ + 16;  // one (copy, channel) plane, padded
    constexpr int GROW = 256 + 16;          // one gradient channel, padded
    constexpr int XONES = 3 * CINP * XCH;   // a plane of ones: the bias row
    constexpr int XZERO = XONES + XCH;      // a plane of zeros: padding rows
    __shared__ __attribute__((aligned(16))) uint16_t xs[3 * CINP * XCH + 2 * XCH];
    __shared__ __attribute__((aligned(16))) uint16_t gs[16 * MT * GROW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mtw = wave % MT;  // pairs p = wave + 4 q all take channel tile p % MT
    const int g4 = lane >> 4, m = lane & 15;
    const int n = a.n, nin = a.up ? n / 2 : n;
    const int nbx = (n + BX - 1) / BX, nby = (n + BY - 1) / BY, nbz = (n + BZ - 1) / BZ;
    const int nbricks = nbx * nby * nbz;
    const v4s* xin = reinterpret_cast<const v4s*>(a.x);

    // per tile pair of this wavefront: LDS element offset of its B row (tap
    // and channel), of the ones plane for the bias row, of the zeros plane
    // for padding rows and for a wavefront's pairs past the last; every
    // fragment is then one unconditional LDS read
    int boff[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) {
        const int p = wave + 4 * q;
        const int r = (p / MT) * 16 + m;
        boff[q] = XZERO;
        if (p < PAIRS && r == NROWS) boff[q] = XONES;
        if (p < PAIRS && r < NROWS) {
            int dx, dy, dz;
            tap_offset(r / CINP, dx, dy, dz);
            boff[q] = ((dz + 1) * CINP + r % CINP) * XCH + ((1 + dx) * LY + (1 + dy)) * BZ;
        }
    }
    v4f acc[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int e = threadIdx.x; e < XCH; e += 256) {  // ordered by the first brick's barrier
        xs[XONES + e] = 0x3F80;                       // 1.0 in bfloat16
        xs[XZERO + e] = 0;
    }

    // Staging, one brick ahead: the loads of brick b + gridDim.x are issued
    // into registers after brick b is in LDS and complete during its MFMAs.
    // Input: positions (x0-1 .. x0+4) x (y0-1 .. y0+4) x (z0-1 .. z0+16),
    // 4-channel groups; position z_in goes to copy s at z = z_in - z0 - s + 1.
    // An input item is one channel group of one (x, y) row and four output z
    // of every copy: the six input voxels they need, written as one 8-byte
    // quad of z per (copy, channel).  A gradient item is one channel group of
    // four consecutive voxels along z, stored transposed (zero outside the
    // grid).
    constexpr int NXI = LX * LY * 4 * G4, PXI = (NXI + 255) / 256;
    constexpr int NGI = 64 * (4 * MT), PGI = (NGI + 255) / 256;
    v4s ux[PXI][6], ug[PGI][4];
    auto load_brick = [&](int b) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * BZ;
#pragma unroll
        for (int it = 0; it < PXI; ++it) {
            const int e = threadIdx.x + 256 * it;
            const int q = e % G4, zq = (e / G4) % 4, vy = (e / (G4 * 4)) % LY,
                      vx = e / (G4 * 4 * LY);
            const int ii = x0 - 1 + vx, jj = y0 - 1 + vy;
            const bool row = e < NXI && ii >= 0 && ii < n && jj >= 0 && jj < n;
            const int si = a.up ? ii >> 1 : ii, sj = a.up ? jj >> 1 : jj;
            const long rbase = ((long)si * nin + sj) * nin;
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const int kz = z0 - 1 + 4 * zq + t;
                ux[it][t] = v4s{0, 0, 0, 0};
                if (row && kz >= 0 && kz < n)
                    ux[it][t] = xin[(rbase + (a.up ? kz >> 1 : kz)) * G4 + q];
            }
        }
#pragma unroll
        for (int it = 0; it < PGI; ++it) {
            const int e = threadIdx.x + 256 * it;
            const int q = e % (4 * MT), v = 4 * (e / (4 * MT));
            const int vz = v % BZ, vy = (v / BZ) % BY, vx = v / (BZ * BY);
            const int i = x0 + vx, j = y0 + vy;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = z0 + vz + t;
                ug[it][t] = v4s{0, 0, 0, 0};
                if (e < NGI && 4 * q < a.coutp && i < n && j < n && k < n)
                    ug[it][t] = *reinterpret_cast<const v4s*>(
                        a.g + (((long)i * n + j) * n + k) * a.coutp + 4 * q);
            }
        }
    };
    if (blockIdx.x < nbricks) load_brick(blockIdx.x);

    for (int b = blockIdx.x; b < nbricks; b += gridDim.x) {
        __syncthreads();  // the previous brick's reads are done
#pragma unroll
        for (int it = 0; it < PXI; ++it) {
            const int e = threadIdx.x + 256 * it;
            if (e >= NXI) continue;
            const int q = e % G4, zq = (e / G4) % 4, vy = (e / (G4 * 4)) % LY,
                      vx = e / (G4 * 4 * LY);
            uint16_t* dst = xs + (4 * q * XCH + (vx * LY + vy) * BZ + 4 * zq);
#pragma unroll
            for (int sft = 0; sft < 3; ++sft)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    *reinterpret_cast<v4s*>(dst + (sft * CINP + c) * XCH) =
                        v4s{ux[it][sft][c], ux[it][sft + 1][c], ux[it][sft + 2][c],
                            ux[it][sft + 3][c]};
        }
#pragma unroll
        for (int it = 0; it < PGI; ++it) {
            const int e = threadIdx.x + 256 * it;
            if (e >= NGI) continue;
            const int q = e % (4 * MT), v = 4 * (e / (4 * MT));
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int co = 4 * q + c;
                *reinterpret_cast<v4s*>(gs + co * GROW + v) =
                    (co < a.cout) ? v4s{ug[it][0][c], ug[it][1][c], ug[it][2][c], ug[it][3][c]}
                                  : v4s{0, 0, 0, 0};
            }
        }
        __syncthreads();
        if (b + (int)gridDim.x < nbricks) load_brick(b + gridDim.x);
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {  // 16 voxels: (vx, vy) = (ks / 4, ks % 4), z 0..15
            const int vbase = ks * 16 + 4 * g4;
            const int xoff = ((ks >> 2) * LY + (ks & 3)) * BZ + 4 * g4;
            const v4s av = *reinterpret_cast<const v4s*>(&gs[(mtw * 16 + m) * GROW + vbase]);
#pragma unroll
            for (int q = 0; q < PW; ++q) {
                const v4s bv = *reinterpret_cast<const v4s*>(&xs[boff[q] + xoff]);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[q], 0, 0, 0);
            }
        }
    }
    // D[co][row] of each pair: lane holds rows nt*16 + m, channels mt*16 + 4 g4 + i
    const long len = (long)NROWS * a.cout + a.cout;
    float* part = a.part + (long)blockIdx.x * len;
#pragma unroll
    for (int q = 0; q < PW; ++q) {
        const int p = wave + 4 * q;
        if (p >= PAIRS) break;
        const int nt = p / MT, mt = p % MT;
        const int r = nt * 16 + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = mt * 16 + 4 * g4 + i;
            if (co >= a.cout) continue;
            if (r < NROWS)
                part[(long)r * a.cout + co] = acc[q][i];
            else if (r == NROWS)
                part[(long)NROWS * a.cout + co] = acc[q][i];
        }
    }
}

// out[q] += sum over blocks b of part[b][q]; q < nw goes to dw, the rest to
// db.  A block takes 64 consecutive q; its 16 rows of 64 threads each sum
// the blocks b = row, row + 16, ... in order, and the 16 row sums are added
// in row order, so the result is fixed for a given block count.
__global__ void __launch_bounds__(1024) wgrad_reduce_kernel(const float* __restrict__ part,
                                                            int nblk, long len, long nw,
                                                            float* __restrict__ dw,
                                                            float* __restrict__ db) {
    __shared__ float s_row[16][64];
    const int c = threadIdx.x & 63, rw = threadIdx.x >> 6;
    const long q = (long)blockIdx.x * 64 + c;
    float s = 0.f;
    if (q < len)
        for (int b = rw; b < nblk; b += 16) s += part[(long)b * len + q];
    s_row[rw][c] = s;
    __syncthreads();
    if (rw == 0 && q < len) {
        float t = s_row[0][c];
#pragma unroll
        for (int r = 1; r < 16; ++r) t += s_row[r][c];
        if (q < nw)
            dw[q] += t;
        else
            db[q - nw] += t;
    }
}

template <int CINP>
hipError_t launch_wgrad_cin(const WgradArgs& a, int nblk, hipStream_t st) {
    if (a.cout > 16)
        hipLaunchKernelGGL((conv_wgrad_lds_kernel<CINP, 2>), dim3(nblk), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((conv_wgrad_lds_kernel<CINP, 1>), dim3(nblk), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t conv_wgrad(const WgradArgs& a, hipStream_t st) {
    if (a.cout < 1 || a.cout > 32 || a.coutp < a.cout || a.coutp % 4 || a.n < 1 ||
        (a.up && a.n % 2) || !a.part)
        return hipErrorInvalidValue;
    const long nbricks = (long)((a.n + BX - 1) / BX) * ((a.n + BY - 1) / BY) * ((a.n + BZ - 1) / BZ);
    const int nblk = (int)(nbricks < WGRAD_MAX_BLOCKS ? nbricks : WGRAD_MAX_BLOCKS);
    const long len = wgrad_part_len(a.cinp, a.cout);
    if (len > wgrad_part_len(24, 24)) return hipErrorInvalidValue;
    hipError_t e;
    switch (a.cinp) {
        case 4: e = launch_wgrad_cin<4>(a, nblk, st); break;
        case 8: e = launch_wgrad_cin<8>(a, nblk, st); break;
        case 12: e = launch_wgrad_cin<12>(a, nblk, st); break;
        case 16: e = launch_wgrad_cin<16>(a, nblk, st); break;
        case 24: e = launch_wgrad_cin<24>(a, nblk, st); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    const long nw = 27L * a.cinp * a.cout;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((len + 63) / 64)), dim3(1024), 0, st,
                       a.part, nblk, len, nw, a.dw, a.db);
    return hipGetLastError();
}

// ---------------------------------------------------------------- packing --
// w: FP32 [27][cinp][cout] (padded input channels stay zero).  Fragment
// (ntile, kstep, lane, i) holds B[k][col] with k = 16 kstep + 4 (lane/16) + i,
// col = 16 ntile + lane % 16.  The adjoint (transposed = 1) is the
// convolution mapping output-channel gradients to input-channel gradients:
// taps mirrored (t -> 26 - t), input and output channels swapped.
__global__ void pack_kernel(const float* __restrict__ w, int cin, int cout, int transposed,
                            uint16_t* __restrict__ wp) {
    const int cinp = (cin + 3) & ~3;
    const int kin = transposed ? cout : cin;            // channels read by the conv
    const int kinp = (kin + 3) & ~3;
    const int kout = transposed ? cin : cout;           // channels produced
    const int ksteps = (27 * kinp + 15) / 16;
    const int ntiles = (kout + 15) / 16;
    const long total = (long)ntiles * ksteps * 64 * 4;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e & 3);
        const int lane = (int)((e >> 2) & 63);
        const long r = e >> 8;
        const int ks = (int)(r % ksteps), nt = (int)(r / ksteps);
        const int k = ks * 16 + 4 * (lane >> 4) + i;
        const int col = nt * 16 + (lane & 15);
        const int t = k / kinp, c = k % kinp;
        float val = 0.f;
        if (t < 27 && c < kin && col < kout) {
            if (!transposed)
                val = w[((long)t * cinp + c) * cout + col];
            else
                val = w[((long)(26 - t) * cinp + col) * cout + c];
        }
        wp[e] = f2bf(val);
    }
}

hipError_t pack_weights(const float* w, int cin, int cout, int transposed, uint16_t* wp,
                        hipStream_t st) {
    hipLaunchKernelGGL(pack_kernel, dim3(64), dim3(256), 0, st, w, cin, cout, transposed, wp);
    return hipGetLastError();
}

// ------------------------------------------------------------ elementwise --
__global__ void avgpool2_kernel(const uint16_t* __restrict__ x, int n, int cp,
                                uint16_t* __restrict__ y) {
    const int nc = n / 2;
    const long total = (long)nc * nc * nc * cp;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e % cp);
        const long v = e / cp;
        const int k = (int)(v % nc), j = (int)((v / nc) % nc), i = (int)(v / ((long)nc * nc));
        float s = 0.f;
        for (int d = 0; d < 8; ++d) {
            const long fi = 2 * i + (d >> 2), fj = 2 * j + ((d >> 1) & 1), fk = 2 * k + (d & 1);
            s += bf2f(x[((fi * n + fj) * n + fk) * cp + c]);
        }
        y[e] = f2bf(0.125f * s);
    }
}

__global__ void sumpool2_kernel(const uint16_t* __restrict__ g, int n, int cp,
                                uint16_t* __restrict__ out, int accumulate) {
    const int nc = n / 2;
    const long total = (long)nc * nc * nc * cp;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e % cp);
        const long v = e / cp;
        const int k = (int)(v % nc), j = (int)((v / nc) % nc), i = (int)(v / ((long)nc * nc));
        float s = accumulate ? bf2f(out[e]) : 0.f;
        for (int d = 0; d < 8; ++d) {
            const long fi = 2 * i + (d >> 2), fj = 2 * j + ((d >> 1) & 1), fk = 2 * k + (d & 1);
            s += bf2f(g[((fi * n + fj) * n + fk) * cp + c]);
        }
        out[e] = f2bf(s);
    }
}

__global__ void unpool_add_kernel(const uint16_t* __restrict__ g, int n, int cp,
                                  uint16_t* __restrict__ out) {
    const int nc = n / 2;
    const long total = (long)n * n * n * cp;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e % cp);
        const long v = e / cp;
        const int k = (int)(v % n), j = (int)((v / n) % n), i = (int)(v / ((long)n * n));
        const long cv = (((long)(i >> 1) * nc + (j >> 1)) * nc + (k >> 1)) * cp + c;
        out[e] = f2bf(bf2f(out[e]) + 0.125f * bf2f(g[cv]));
    }
}

__global__ void relu_mask_kernel(const uint16_t* __restrict__ g, const uint16_t* __restrict__ r,
                                 long count, uint16_t* __restrict__ out) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        out[e] = (bf2f(r[e]) > 0.f) ? g[e] : (uint16_t)0;
}

__global__ void add_kernel(uint16_t* __restrict__ a, const uint16_t* __restrict__ b, long count) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        a[e] = f2bf(bf2f(a[e]) + bf2f(b[e]));
}

__global__ void copy_kernel(const uint16_t* __restrict__ a, long count, uint16_t* __restrict__ o) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        o[e] = a[e];
}

// sums over a block: wave shuffles, then one atomic per block
__device__ __forceinline__ void block_add2(float s, float s2, float* acc) {
    __shared__ float part[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = s;
        part[1][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.f, b = 0.f;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            a += part[0][i];
            b += part[1][i];
        }
        atomicAdd(&acc[0], a);
        atomicAdd(&acc[1], b);
    }
}

template <class T>
__global__ void __launch_bounds__(256) log_stats_kernel(const T* __restrict__ rho, long nvox,
                                                        float* acc) {
    float s = 0.f, s2 = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long)gridDim.x * blockDim.x) {
        const float l = log10f(fmaxf((float)rho[v], 1e-12f));
        s += l;
        s2 += l * l;
    }
    block_add2(s, s2, acc);
}

template <class T>
__global__ void make_features_kernel(const T* __restrict__ rho, int n,
                                     const float* __restrict__ acc,
                                     const double* __restrict__ mom, double x0, double y0,
                                     double z0, double L, uint16_t* __restrict__ x,
                                     float* __restrict__ mono) {
    const long nvox = (long)n * n * n;
    const float mu = acc[0] / (float)nvox;
    const float var = fmaxf(acc[1] / (float)nvox - mu * mu, 1e-12f);
    const float inv = 1.0f / sqrtf(var);
    // centre of mass in box units
    const float cx = (float)((mom[1] - x0) / L), cy = (float)((mom[2] - y0) / L),
                cz = (float)((mom[3] - z0) / L);
    const float h = 1.0f / (float)n, eps2 = (1.5f * h) * (1.5f * h);
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % n), j = (int)((v / n) % n), i = (int)(v / ((long)n * n));
        const float l = log10f(fmaxf((float)rho[v], 1e-12f));
        const float dx = (i + 0.5f) * h - cx, dy = (j + 0.5f) * h - cy, dz = (k + 0.5f) * h - cz;
        const float m = -1.0f / sqrtf(dx * dx + dy * dy + dz * dz + eps2);
        mono[v] = m;
        x[4 * v] = f2bf((l - mu) * inv);
        x[4 * v + 1] = f2bf(0.1f * m);
        x[4 * v + 2] = 0;
        x[4 * v + 3] = 0;
    }
}

__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ p,
                                                  const float* __restrict__ t, long nvox,
                                                  float scale, uint16_t* __restrict__ g,
                                                  float* loss) {
    float s = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long)gridDim.x * blockDim.x) {
        const float d = p[v] - t[v];
        s += d * d;
        g[4 * v] = f2bf(2.0f * d * scale);
        g[4 * v + 1] = 0;
        g[4 * v + 2] = 0;
        g[4 * v + 3] = 0;
    }
    __shared__ float part[4];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) * scale);
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, long count, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x) {
        const float gr = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gr;
        const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
        m[e] = mm;
        v[e] = vv;
        w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

__global__ void restrict2_kernel(const double* __restrict__ f, int n, double* __restrict__ c) {
    const int nc = n / 2;
    const long total = (long)nc * nc * nc;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e % nc), j = (int)((e / nc) % nc), i = (int)(e / ((long)nc * nc));
        double s = 0.0;
        for (int d = 0; d < 8; ++d) {
            const long fi = 2 * i + (d >> 2), fj = 2 * j + ((d >> 1) & 1), fk = 2 * k + (d & 1);
            s += f[(fi * n + fj) * n + fk];
        }
        c[e] = 0.125 * s;
    }
}

// coarse index pair and weight of fine cell i (fine grid n = f nc): the
// fine centre sits at coarse coordinate (i + 1/2) / f - 1/2
__device__ __forceinline__ void lin_weights(int i, int f, int nc, int& i0, int& i1, float& w1) {
    const float xc = (i + 0.5f) / (float)f - 0.5f;
    int a = (int)floorf(xc);
    float w = xc - (float)a;
    if (a < 0) { a = 0; w = 0.0f; }
    if (a >= nc - 1) { a = nc - 1; w = 0.0f; }
    i0 = a;
    i1 = min(a + 1, nc - 1);
    w1 = w;
}

__global__ void prolong_kernel(const float* __restrict__ c, int nc, int n,
                               const double* __restrict__ mom, double G, double L,
                               double* __restrict__ fine) {
    const long total = (long)n * n * n;
    const int f = n / nc;
    const double scale = G * mom[0] / L;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e % n), j = (int)((e / n) % n), i = (int)(e / ((long)n * n));
        int i0, i1, j0, j1, k0, k1;
        float wi, wj, wk;
        lin_weights(i, f, nc, i0, i1, wi);
        lin_weights(j, f, nc, j0, j1, wj);
        lin_weights(k, f, nc, k0, k1, wk);
        auto at = [&](int a, int b, int d) { return c[((long)a * nc + b) * nc + d]; };
        const float c00 = at(i0, j0, k0) * (1 - wk) + at(i0, j0, k1) * wk;
        const float c01 = at(i0, j1, k0) * (1 - wk) + at(i0, j1, k1) * wk;
        const float c10 = at(i1, j0, k0) * (1 - wk) + at(i1, j0, k1) * wk;
        const float c11 = at(i1, j1, k0) * (1 - wk) + at(i1, j1, k1) * wk;
        const float c0 = c00 * (1 - wj) + c01 * wj, c1 = c10 * (1 - wj) + c11 * wj;
        fine[e] = scale * (double)(c0 * (1 - wi) + c1 * wi);
    }
}

// ------------------------------------------------- start correction --
// The correction network reads the warm start's residual averaged over
// f^3 blocks (nc = n / f points per direction) and predicts the correction
// there; the correction is interpolated back and added to the start.

// r (FP32, red-black split layout of the n^3 grid, n/2 entries per row and
// colour) averaged over f^3 blocks into a natural nc^3 array; each block is
// summed in a fixed order.
__global__ void corr_restrict_kernel(const float* __restrict__ rr, const float* __restrict__ rb,
                                     int n, int f, float* __restrict__ out) {
    const int nc = n / f, nzh = n / 2;
    const long total = (long)nc * nc * nc;
    const float w = 1.0f / (float)(f * f * f);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int K = (int)(e % nc), J = (int)((e / nc) % nc), I = (int)(e / ((long)nc * nc));
        float s = 0.0f;
        for (int a = 0; a < f; ++a)
            for (int b = 0; b < f; ++b)
                for (int c = 0; c < f; ++c) {
                    const int i = f * I + a, j = f * J + b, k = f * K + c;
                    const float* src = ((i + j + k) & 1) ? rb : rr;
                    s += src[((long)i * n + j) * nzh + (k >> 1)];
                }
        out[e] = s * w;
    }
}

// Sum of x^2 over count values in a fixed order: CORR_SUM_BLOCKS
// workgroups of 256 each sum a strided share and reduce it by a tree
// (first kernel, x != nullptr), then one workgroup adds the partial sums by
// the same tree (second kernel, x == nullptr).
constexpr int CORR_SUM_BLOCKS = 256;
__global__ void __launch_bounds__(256) corr_sumsq_kernel(const float* __restrict__ x, long count,
                                                         double* __restrict__ part,
                                                         double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    if (x) {
        const long stride = (long)CORR_SUM_BLOCKS * 256;
        for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < count; v += stride) {
            const double a = (double)x[v];
            s += a * a;
        }
    } else {
        s = part[threadIdx.x];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) (x ? part[blockIdx.x] : out[0]) = red[0];
}

// Features of the correction network (cp = 4 layout): channel 0 = r / s,
// channel 1 = -r / s, s the root mean square of r (r all zero: zeros); no
// monopole term.
__global__ void corr_features_kernel(const float* __restrict__ r, long nvox,
                                     const double* __restrict__ sumsq, uint16_t* __restrict__ x,
                                     float* __restrict__ mono) {
    const double ms = sumsq[0] / (double)nvox;
    const float inv = ms > 0.0 ? (float)(1.0 / sqrt(ms)) : 0.0f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long)gridDim.x * blockDim.x) {
        const float a = r[v] * inv;
        x[4 * v] = f2bf(a);
        x[4 * v + 1] = f2bf(-a);
        x[4 * v + 2] = 0;
        x[4 * v + 3] = 0;
        mono[v] = 0.0f;
    }
}

// Catmull-Rom taps of fine cell i (f fine cells per coarse cell, nc coarse
// cells, cell centres): the coarse cells a - 1 .. a + 2 around the coarse
// coordinate x = (i + 0.5) / f - 0.5 with the cubic convolution weights of
// t = x - a; a tap outside the grid takes the mirrored cell with its weight
// negated (an odd reflection: the correction vanishes at the boundary).
// Trilinear interpolation leaves a Laplacian spike at every coarse cell
// centre (the slope jumps there), which makes the interpolated correction's
// residual larger than the one it removes once f reaches 4; this cubic is
// continuously differentiable.
__device__ __forceinline__ void cr_taps(int i, int f, int nc, int idx[4], float w[4]) {
    const float x = ((float)i + 0.5f) / (float)f - 0.5f;
    const float af = floorf(x);
    const float t = x - af, t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t3 + 2.0f * t2 - t);
    w[1] = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
    w[2] = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
    w[3] = 0.5f * (t3 - t2);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int c = (int)af - 1 + q;
        if (c < 0) {
            c = -c - 1;
            w[q] = -w[q];
        } else if (c >= nc) {
            c = 2 * nc - c - 1;
            w[q] = -w[q];
        }
        idx[q] = c;
    }
}

// Separable interpolation of e (nc^3, natural) to the fine grid, first
// along z: t1[I][J][k] (nc x nc x n), then along y: t2[I][j][k] (nc x n x n).
__global__ void corr_interp_z_kernel(const float* __restrict__ e, int nc, int n,
                                     float* __restrict__ t1) {
    const long total = (long)nc * nc * n;
    const int f = n / nc;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < total;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % n);
        const long row = v / n;  // I * nc + J
        int idx[4];
        float w[4];
        cr_taps(k, f, nc, idx, w);
        const float* src = e + row * nc;
        t1[v] = w[0] * src[idx[0]] + w[1] * src[idx[1]] + w[2] * src[idx[2]] + w[3] * src[idx[3]];
    }
}

__global__ void corr_interp_y_kernel(const float* __restrict__ t1, int nc, int n,
                                     float* __restrict__ t2) {
    const long total = (long)nc * n * n;
    const int f = n / nc;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < total;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % n);
        const int j = (int)((v / n) % n);
        const int I = (int)(v / ((long)n * n));
        int idx[4];
        float w[4];
        cr_taps(j, f, nc, idx, w);
        const float* src = t1 + (long)I * nc * n + k;
        t2[v] = w[0] * src[(long)idx[0] * n] + w[1] * src[(long)idx[1] * n] +
                w[2] * src[(long)idx[2] * n] + w[3] * src[(long)idx[3] * n];
    }
}

// The correction scale * s * P(e) at the interior cells of the split n^3
// grid and 0 at its boundary cells, in FP32 (the inner solve's unknown): P
// the separable Catmull-Rom interpolation above, finished here along x from
// t2; s the root mean square of the averaged residual.  One thread per
// split entry (i, j, m), both colours.
__global__ void corr_apply_kernel(const float* __restrict__ t2, int nc, int n,
                                  const double* __restrict__ sumsq, double scale,
                                  float* __restrict__ er, float* __restrict__ eb) {
    const int nzh = n / 2, f = n / nc;
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (m >= nzh || j >= n) return;
    const double sc = scale * sqrt(sumsq[0] / ((double)nc * nc * nc));
    const long base = ((long)i * n + j) * nzh + m;
    int idx[4];
    float w[4];
    cr_taps(i, f, nc, idx, w);
    const bool inner_ij = i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2;
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
        const int k = 2 * m + ((i + j + colour) & 1);
        float v = 0.0f;
        if (inner_ij && k >= 1 && k <= n - 2) {
            const float* src = t2 + (long)j * n + k;
            const long pl = (long)n * n;
            const float c = w[0] * src[idx[0] * pl] + w[1] * src[idx[1] * pl] +
                            w[2] * src[idx[2] * pl] + w[3] * src[idx[3] * pl];
            v = (float)(sc * (double)c);
        }
        (colour ? eb : er)[base] = v;
    }
}

// ------------------------------------------------------------ launchers --
static unsigned grid_n(long n) {
    long b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

hipError_t avgpool2(const uint16_t* x, int n, int cp, uint16_t* y, hipStream_t st) {
    const long tot = (long)(n / 2) * (n / 2) * (n / 2) * cp;
    hipLaunchKernelGGL(avgpool2_kernel, dim3(grid_n(tot)), dim3(256), 0, st, x, n, cp, y);
    return hipGetLastError();
}
hipError_t sumpool2_add(const uint16_t* g, int n, int cp, uint16_t* out, int accumulate,
                        hipStream_t st) {
    const long tot = (long)(n / 2) * (n / 2) * (n / 2) * cp;
    hipLaunchKernelGGL(sumpool2_kernel, dim3(grid_n(tot)), dim3(256), 0, st, g, n, cp, out,
                       accumulate);
    return hipGetLastError();
}
hipError_t unpool_add(const uint16_t* g, int n, int cp, uint16_t* out, hipStream_t st) {
    const long tot = (long)n * n * n * cp;
    hipLaunchKernelGGL(unpool_add_kernel, dim3(grid_n(tot)), dim3(256), 0, st, g, n, cp, out);
    return hipGetLastError();
}
hipError_t relu_mask(const uint16_t* g, const uint16_t* r, long count, uint16_t* out,
                     hipStream_t st) {
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_n(count)), dim3(256), 0, st, g, r, count, out);
    return hipGetLastError();
}
hipError_t add_inplace(uint16_t* a, const uint16_t* b, long count, hipStream_t st) {
    hipLaunchKernelGGL(add_kernel, dim3(grid_n(count)), dim3(256), 0, st, a, b, count);
    return hipGetLastError();
}
hipError_t copy_bf16(const uint16_t* a, long count, uint16_t* out, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid_n(count)), dim3(256), 0, st, a, count, out);
    return hipGetLastError();
}
static unsigned grid_red(long n) {
    const unsigned g = grid_n(n);
    return g > 512 ? 512 : g;
}

hipError_t log_stats(const float* rho, long nvox, float* acc, hipStream_t st) {
    hipLaunchKernelGGL(log_stats_kernel<float>, dim3(grid_red(nvox)), dim3(256), 0, st, rho, nvox,
                       acc);
    return hipGetLastError();
}
hipError_t log_stats_f64(const double* rho, long nvox, float* acc, hipStream_t st) {
    hipLaunchKernelGGL(log_stats_kernel<double>, dim3(grid_red(nvox)), dim3(256), 0, st, rho,
                       nvox, acc);
    return hipGetLastError();
}
hipError_t make_features(const float* rho, int n, const float* acc, const double* mom,
                         double x0, double y0, double z0, double L, uint16_t* x, float* mono,
                         hipStream_t st) {
    const long nv = (long)n * n * n;
    hipLaunchKernelGGL(make_features_kernel<float>, dim3(grid_n(nv)), dim3(256), 0, st, rho, n,
                       acc, mom, x0, y0, z0, L, x, mono);
    return hipGetLastError();
}
hipError_t make_features_f64(const double* rho, int n, const float* acc, const double* mom,
                             double x0, double y0, double z0, double L, uint16_t* x, float* mono,
                             hipStream_t st) {
    const long nv = (long)n * n * n;
    hipLaunchKernelGGL(make_features_kernel<double>, dim3(grid_n(nv)), dim3(256), 0, st, rho, n,
                       acc, mom, x0, y0, z0, L, x, mono);
    return hipGetLastError();
}
hipError_t mse_loss(const float* pred, const float* target, long nvox, float scale, uint16_t* g,
                    float* loss, hipStream_t st) {
    hipLaunchKernelGGL(mse_kernel, dim3(grid_red(nvox)), dim3(256), 0, st, pred, target, nvox,
                       scale, g, loss);
    return hipGetLastError();
}
hipError_t adam_step(float* w, const float* g, float* m, float* v, long count, float lr, float b1,
                     float b2, float eps, float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(grid_n(count)), dim3(256), 0, st, w, g, m, v, count, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t restrict2_f64(const double* fine, int n, double* coarse, hipStream_t st) {
    const long tot = (long)(n / 2) * (n / 2) * (n / 2);
    hipLaunchKernelGGL(restrict2_kernel, dim3(grid_n(tot)), dim3(256), 0, st, fine, n, coarse);
    return hipGetLastError();
}
hipError_t prolong_to_f64(const float* coarse, int nc, int n, const double* mom, double G,
                          double L, double* fine, hipStream_t st) {
    if (nc <= 0 || n % nc) return hipErrorInvalidValue;
    const long tot = (long)n * n * n;
    hipLaunchKernelGGL(prolong_kernel, dim3(grid_n(tot)), dim3(256), 0, st, coarse, nc, n, mom, G,
                       L, fine);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace hyd

// ---------------------------------------------------------- training data --
namespace hyd {
namespace nn {

// One density field: blobs (Gaussian or n=1 polytrope profiles) on a floor,
// times exp(delta - sigma^2/2) with delta a sum of Fourier modes (amplitude
// ~ 1/k, i.e. power ~ k^-2).  blob = [cx cy cz mass width type] x nblob,
// mode = [kx ky kz phase amp] x nmode (integer k: periodic on the box).
__global__ void blob_field_kernel(int n, const double* __restrict__ blob, int nblob,
                                  const double* __restrict__ mode, int nmode, double floor_rho,
                                  double sigma, double* __restrict__ rho64,
                                  float* __restrict__ rho32) {
    const long total = (long)n * n * n;
    const double h = 1.0 / n;
    const double twopi = 6.283185307179586;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < total;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % n), j = (int)((v / n) % n), i = (int)(v / ((long)n * n));
        const double x = (i + 0.5) * h, y = (j + 0.5) * h, z = (k + 0.5) * h;
        double r = floor_rho;
        for (int b = 0; b < nblob; ++b) {
            const double* p = blob + 6 * b;
            const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
            const double rr = sqrt(dx * dx + dy * dy + dz * dz);
            const double w = p[4];
            if (p[5] < 0.5) {
                const double norm = 15.749609945653303 * w * w * w;  // (2 pi)^1.5 w^3
                r += p[3] * exp(-0.5 * (rr / w) * (rr / w)) / norm;
            } else {
                const double R = 2.0 * w;
                if (rr < R) {
                    const double q = 3.141592653589793 * rr / R;
                    const double prof = q > 0.0 ? sin(q) / q : 1.0;
                    r += p[3] * prof / (4.0 * R * R * R / 3.141592653589793);
                }
            }
        }
        double d = 0.0;
        for (int m = 0; m < nmode; ++m) {
            const double* p = mode + 5 * m;
            d += p[4] * cos(twopi * (p[0] * x + p[1] * y + p[2] * z) + p[3]);
        }
        r *= exp(d - 0.5 * sigma * sigma);
        rho64[v] = r;
        if (rho32) rho32[v] = (float)r;
    }
}

hipError_t blob_field(int n, const double* blob, int nblob, const double* mode, int nmode,
                      double floor_rho, double sigma, double* rho64, float* rho32,
                      hipStream_t st) {
    const long tot = (long)n * n * n;
    hipLaunchKernelGGL(blob_field_kernel, dim3(grid_n(tot)), dim3(256), 0, st, n, blob, nblob,
                       mode, nmode, floor_rho, sigma, rho64, rho32);
    return hipGetLastError();
}

// target potential in units of G M / L, natural layout, FP32
__global__ void scale_f64_to_f32_kernel(const double* __restrict__ a, long count,
                                        const double* __restrict__ mom, double G, double L,
                                        float* __restrict__ out) {
    const double s = (mom[0] > 0.0) ? L / (G * mom[0]) : 1.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        out[e] = (float)(a[e] * s);
}

hipError_t scale_potential(const double* phi, long count, const double* mom, double G, double L,
                           float* out, hipStream_t st) {
    hipLaunchKernelGGL(scale_f64_to_f32_kernel, dim3(grid_n(count)), dim3(256), 0, st, phi, count,
                       mom, G, L, out);
    return hipGetLastError();
}

hipError_t corr_restrict(const float* rr, const float* rb, int n, int f, float* out,
                         hipStream_t st) {
    const long total = (long)(n / f) * (n / f) * (n / f);
    hipLaunchKernelGGL(corr_restrict_kernel, dim3(grid_n(total)), dim3(256), 0, st, rr, rb, n, f,
                       out);
    return hipGetLastError();
}
hipError_t corr_features(const float* r, int nc, double* sumsq, double* part, uint16_t* x,
                         float* mono, hipStream_t st) {
    const long nv = (long)nc * nc * nc;
    hipLaunchKernelGGL(corr_sumsq_kernel, dim3(CORR_SUM_BLOCKS), dim3(256), 0, st, r, nv, part,
                       sumsq);
    hipLaunchKernelGGL(corr_sumsq_kernel, dim3(1), dim3(256), 0, st, (const float*)nullptr, nv,
                       part, sumsq);
    hipLaunchKernelGGL(corr_features_kernel, dim3(grid_n(nv)), dim3(256), 0, st, r, nv, sumsq, x,
                       mono);
    return hipGetLastError();
}
hipError_t corr_apply(const float* e, int nc, int n, float* t1, float* t2, const double* sumsq,
                      double scale, float* er, float* eb, hipStream_t st) {
    hipLaunchKernelGGL(corr_interp_z_kernel, dim3(grid_n((long)nc * nc * n)), dim3(256), 0, st, e,
                       nc, n, t1);
    hipLaunchKernelGGL(corr_interp_y_kernel, dim3(grid_n((long)nc * n * n)), dim3(256), 0, st, t1,
                       nc, n, t2);
    hipLaunchKernelGGL(corr_apply_kernel,
                       dim3((unsigned)((n / 2 + 63) / 64), (unsigned)((n + 3) / 4), (unsigned)n),
                       dim3(256), 0, st, t2, nc, n, sumsq, scale, er, eb);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace hyd
