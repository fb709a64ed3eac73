// seqsum.hip -- sequential-order sums (see seqsum.hpp for the method).
//
// Composite method, per sequence q of npad terms:
//   seq_blocksum_kernel  approximate sum of every 1024-term block
//   seq_scan_kernel      approximate exclusive prefix of those sums
//   seq_prep_kernel      per chunk: predicted binade, D, lo, hi; per run of
//                        chunks sharing a binade (a segment), the same
//                        composed; per block of 1024: its segment when one
//                        spans it
//   seq_level2_kernel    the same for blocks of 16384
//   seq_walk_kernel      one wavefront moves the running sum through the
//                        blocks, applying a block's increment when its
//                        conditions hold and adding the terms of a chunk one
//                        by one when they do not
// The approximate sums only choose which binade to try; every application
// is checked against the exact running sum, so a wrong guess costs time, not
// bits.
#include "seqsum.hpp"

#include <climits>

namespace hyd {

namespace {

constexpr long long SEQ_LIM = 1LL << 55;  // any |D|, lo, hi beyond this never applies
constexpr long long SEQ_2P52 = 1LL << 52;
constexpr long long SEQ_2P53 = 1LL << 53;
constexpr int SEQ_CHAIN_TILE = 2048;
constexpr int SEQ_CHAIN_ROW = SEQ_CHAIN_TILE + 32;  // a tile and the overread of its last reads

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ long long rl64(long long v, int lane) {
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = (unsigned)rl((int)(unsigned)(u & 0xffffffffull), lane);
    const unsigned hi = (unsigned)rl((int)(unsigned)(u >> 32), lane);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ SeqDesc rldesc(const SeqDesc& d, int lane) {
    SeqDesc r;
    r.D = rl64(d.D, lane);
    r.lo = rl64(d.lo, lane);
    r.hi = rl64(d.hi, lane);
    r.k = rl(d.k, lane);
    r.flags = rl(d.flags, lane);
    return r;
}

// Move the running sum (its bits sb) over a block described by d when that
// is exact; false leaves it.  Works on the bits: s in the block's binade k
// has exponent field k + 1023 and S = 2^52 + mantissa, or is 2^(k+1) itself
// (S = 2^53); S + D lies in [2^52, 2^53] when the checks pass.  Only
// integer arithmetic on values every lane shares, so it stays in scalar
// registers.
__device__ __forceinline__ bool seq_apply(const SeqDesc& d, long long& sb) {
    if (d.flags & SEQ_ZERO) return true;  // s + (+-0) = s for every s but -0, which never occurs
    if (!(d.flags & SEQ_VALID)) return false;
    const unsigned long long u = (unsigned long long)sb;
    const unsigned long long sign = u >> 63;
    if (sign != (unsigned long long)((d.flags & SEQ_NEG) ? 1 : 0)) return false;
    const int ex = (int)((u >> 52) & 0x7ff);
    const long long man = (long long)(u & ((1ull << 52) - 1));
    long long S;
    if (ex == d.k + 1023) {
        S = man | SEQ_2P52;
    } else if (ex == d.k + 1024 && man == 0) {
        S = SEQ_2P53;
    } else {
        return false;
    }
    if (S + d.lo < SEQ_2P52 || S + d.hi >= SEQ_2P53) return false;
    const long long T = S + d.D;
    const unsigned long long r =
        T == SEQ_2P53 ? (unsigned long long)(d.k + 1024) << 52
                      : ((unsigned long long)(d.k + 1023) << 52) | (unsigned long long)(T - SEQ_2P52);
    sb = (long long)(r | (sign << 63));
    return true;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ long long wave_min64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ long long wave_max64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// One wavefront adds the npad terms at x in order; the other three of the
// 256-thread workgroup stage the next tile in LDS.  Returns the sum (valid in
// every thread of the first wavefront).
__device__ __forceinline__ double seq_chain_group(const double* __restrict__ x, long npad,
                                                  double (*buf)[SEQ_CHAIN_ROW]) {
    const long ntiles = npad / SEQ_CHAIN_TILE;
    const int tid = threadIdx.x;
    for (int e = tid; e < SEQ_CHAIN_TILE; e += 256) buf[0][e] = x[e];
    __syncthreads();
    double s = 0.0;
    for (long t = 0; t < ntiles; ++t) {
        if (tid >= 64) {
            if (t + 1 < ntiles) {
                const double2* src = reinterpret_cast<const double2*>(x + (t + 1) * SEQ_CHAIN_TILE);
                double2* dst = reinterpret_cast<double2*>(buf[(t + 1) & 1]);
                double2 v[6];
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    const int e = tid - 64 + 192 * u;
                    if (e < SEQ_CHAIN_TILE / 2) v[u] = src[e];
                }
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    const int e = tid - 64 + 192 * u;
                    if (e < SEQ_CHAIN_TILE / 2) dst[e] = v[u];
                }
            }
        } else {
            // the next 16 terms are read while the current 16 are added
            const double2* b = reinterpret_cast<const double2*>(buf[t & 1]);
            double2 a[8], c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = b[u];
            for (int e = 0; e < SEQ_CHAIN_TILE / 2; e += 16) {
#pragma unroll
                for (int u = 0; u < 8; ++u) c[u] = b[e + 8 + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s = s + a[u].x;
                    s = s + a[u].y;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = b[e + 16 + u];  // past the tile: padding, unused
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s = s + c[u].x;
                    s = s + c[u].y;
                }
            }
        }
        __syncthreads();
    }
    return s;
}

__global__ void __launch_bounds__(256) seq_chain_kernel(const double* __restrict__ terms,
                                                        long npad, double* __restrict__ sums) {
    __shared__ double buf[2][SEQ_CHAIN_ROW];
    const double s = seq_chain_group(terms + (long)blockIdx.x * npad, npad, buf);
    if (threadIdx.x == 0) {
        sums[2 * blockIdx.x] = s;
        sums[2 * blockIdx.x + 1] = 0.0;
    }
}

__device__ __forceinline__ void load_chunk(const double* __restrict__ p, double (&v)[SEQ_CHUNK]) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK / 2; ++e) {
        const double2 w = p2[e];
        v[2 * e] = w.x;
        v[2 * e + 1] = w.y;
    }
}

// Approximate sum of each 1024-term block (wavefront per block).
__global__ void __launch_bounds__(256) seq_blocksum_kernel(const double* __restrict__ terms,
                                                           long npad, double* __restrict__ bs) {
    const long n1 = npad / SEQ_BLOCK1;
    const long b1 = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b1 >= n1) return;
    const int q = blockIdx.y;
    double v[SEQ_CHUNK];
    load_chunk(terms + (long)q * npad + b1 * SEQ_BLOCK1 + lane * SEQ_CHUNK, v);
    double c = 0.0;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) c += v[e];
    c = wave_sum(c);
    if (lane == 0) bs[(long)q * n1 + b1] = c;
}

// Exclusive prefix of the n1 block sums of sequence blockIdx.x (approximate).
__global__ void __launch_bounds__(1024) seq_scan_kernel(const double* __restrict__ bs, long n1,
                                                        double* __restrict__ pre) {
    __shared__ double s_tot[1024];
    const int tid = threadIdx.x;
    const long per = (n1 + 1023) / 1024;
    const long b = (long)tid * per;
    const long e = b + per < n1 ? b + per : n1;
    const double* in = bs + (long)blockIdx.x * n1;
    double* out = pre + (long)blockIdx.x * n1;
    double t = 0.0;
    for (long i = b; i < e; ++i) t += in[i];
    s_tot[tid] = t;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const double add = tid >= o ? s_tot[tid - o] : 0.0;
        __syncthreads();
        s_tot[tid] += add;
        __syncthreads();
    }
    double run = s_tot[tid] - t;
    for (long i = b; i < e; ++i) {
        out[i] = run;
        run += in[i];
    }
}

// Chunk and 1024-block descriptors (wavefront per block, lane per chunk).
__global__ void __launch_bounds__(256) seq_prep_kernel(const double* __restrict__ terms, long npad,
                                                       const double* __restrict__ pre,
                                                       SeqDesc* __restrict__ dseg,
                                                       SeqDesc* __restrict__ d1) {
    const long n1 = npad / SEQ_BLOCK1;
    const long b1 = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b1 >= n1) return;
    const int q = blockIdx.y;
    double v[SEQ_CHUNK];
    load_chunk(terms + (long)q * npad + b1 * SEQ_BLOCK1 + lane * SEQ_CHUNK, v);
    double c = 0.0;
    bool zero = true;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        c += v[e];
        zero = zero && v[e] == 0.0;
    }
    // approximate running sum at the chunk's start: its binade is the guess
    double incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double w = __shfl_up(incl, o, 64);
        if (lane >= o) incl += w;
    }
    const double P = pre[(long)q * n1 + b1] + (incl - c);
    int ex = 0;
    (void)frexp(P, &ex);
    const int k = ex - 1;  // |P| in [2^k, 2^(k+1))
    const bool neg = P < 0.0;
    bool ok = P != 0.0 && P - P == 0.0 && k >= -960 && k <= 1000;
    long long E = 0, lo = SEQ_LIM, hi = -SEQ_LIM;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        const double sx = neg ? -v[e] : v[e];
        const double t = ldexp(sx, 52 - k);  // exact unless subnormal (then |t| < 0.5)
        const double at = fabs(t);
        long long f = 0, d = 0;
        if (!(at < 0x1p54)) {
            ok = false;
        } else if (at < 0.5) {
            f = sx < 0.0 ? -1 : 0;
        } else {
            const double fl = floor(t);
            if (t - fl == 0.5) ok = false;  // a tie: rounds by the parity of S
            f = (long long)fl;
            d = (long long)rint(t);
        }
        lo = E + f < lo ? E + f : lo;
        hi = E + f > hi ? E + f : hi;
        E += d;
    }
    if (E >= SEQ_LIM || E <= -SEQ_LIM || lo <= -SEQ_LIM || hi >= SEQ_LIM) ok = false;
    if (!ok || zero) E = lo = hi = 0;
    // Segments: maximal runs of chunks that share a binade and sign, every
    // non-zero one valid (zero chunks join the run they fall in; an invalid
    // chunk ends a run and forms its own).  Lane l's segment entry (kept at
    // the lane where a segment starts) composes the chunks [l, end) and
    // carries end in flags >> 8.  The 1024-block is the segment at lane 0
    // when it spans all 64 chunks.
    const bool nz = !zero;
    const int code = k * 2 + (neg ? 1 : 0);
    int lastnz = nz ? lane : -1;  // last non-zero lane <= l
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(lastnz, o, 64);
        if (lane >= o) lastnz = w > lastnz ? w : lastnz;
    }
    int prev = __shfl_up(lastnz, 1, 64);  // last non-zero lane < l
    if (lane == 0) prev = -1;
    const int okp = __shfl(ok ? 1 : 0, prev < 0 ? 0 : prev, 64);
    const int codep = __shfl(code, prev < 0 ? 0 : prev, 64);
    const bool start = lane == 0 || (nz && prev >= 0 && (!ok || !okp || code != codep));
    const unsigned long long starts = __ballot(start);
    // end: the next start after l (64 when none)
    const unsigned long long after = lane == 63 ? 0ull : (starts >> (lane + 1)) << (lane + 1);
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    // first lane of l's segment
    const unsigned long long upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
    const int first = 63 - __clzll((long long)upto);
    long long inc = nz ? E : 0;
    const long long own = inc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long w = __shfl_up(inc, o, 64);
        if (lane >= o) inc += w;
    }
    const long long excl = inc - own;
    const long long base = __shfl(excl, first, 64);  // increments before the segment
    const long long rel = excl - base;
    long long slo = nz ? rel + lo : SEQ_LIM, shi = nz ? rel + hi : -SEQ_LIM;
    int cmin = nz ? code : INT_MAX, cmax = nz ? code : INT_MIN, allok = (!nz || ok) ? 1 : 0;
    // suffix reductions inside the segment: lane l folds in lane l + o when
    // that lane is still in its segment
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long wl = __shfl_down(slo, o, 64), wh = __shfl_down(shi, o, 64);
        const int a = __shfl_down(cmin, o, 64), bb = __shfl_down(cmax, o, 64);
        const int w = __shfl_down(allok, o, 64);
        if (lane + o < end) {
            slo = wl < slo ? wl : slo;
            shi = wh > shi ? wh : shi;
            cmin = a < cmin ? a : cmin;
            cmax = bb > cmax ? bb : cmax;
            allok &= w;
        }
    }
    const long long Dseg = __shfl(inc, end - 1, 64) - base;
    SeqDesc sg;
    if (cmin == INT_MAX) {  // only zero chunks
        sg.D = sg.lo = sg.hi = 0;
        sg.k = 0;
        sg.flags = SEQ_ZERO;
    } else {
        const bool v = allok && cmin == cmax && Dseg < SEQ_LIM && Dseg > -SEQ_LIM &&
                       slo > -SEQ_LIM && shi < SEQ_LIM;
        sg.D = v ? Dseg : 0;
        sg.lo = v ? slo : 0;
        sg.hi = v ? shi : 0;
        sg.k = cmin >> 1;  // floor division keeps k for negative codes
        sg.flags = (v ? SEQ_VALID : 0) | ((cmin & 1) ? SEQ_NEG : 0);
    }
    sg.flags |= (end << 8) | (start ? SEQ_START : 0);
    dseg[((long)q * n1 + b1) * 64 + lane] = sg;
    if (lane == 0) {
        SeqDesc b = sg;
        if (end != 64) {
            b.D = b.lo = b.hi = 0;
            b.flags = 0;
        }
        b.flags &= SEQ_VALID | SEQ_NEG | SEQ_ZERO;
        d1[(long)q * n1 + b1] = b;
    }
}

// 16384-blocks from their sixteen 1024-blocks (thread per block).
__global__ void __launch_bounds__(256) seq_level2_kernel(const SeqDesc* __restrict__ d1, long n2,
                                                         int nq, SeqDesc* __restrict__ d2) {
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= n2 * nq) return;
    const SeqDesc* c = d1 + id * 16;
    bool valid = true, any = false;
    int k = 0, flags = 0;
    long long E = 0, lo = SEQ_LIM, hi = -SEQ_LIM;
    for (int j = 0; j < 16; ++j) {
        const SeqDesc d = c[j];
        if (d.flags & SEQ_ZERO) continue;
        if (!(d.flags & SEQ_VALID)) valid = false;
        if (!any) {
            k = d.k;
            flags = d.flags & SEQ_NEG;
            any = true;
        } else if (d.k != k || (d.flags & SEQ_NEG) != flags) {
            valid = false;
        }
        if (!valid) break;
        lo = E + d.lo < lo ? E + d.lo : lo;
        hi = E + d.hi > hi ? E + d.hi : hi;
        E += d.D;
        if (E >= SEQ_LIM || E <= -SEQ_LIM || lo <= -SEQ_LIM || hi >= SEQ_LIM) valid = false;
    }
    SeqDesc r;
    if (!any) {
        r.D = r.lo = r.hi = 0;
        r.k = 0;
        r.flags = SEQ_ZERO;
    } else {
        r.D = valid ? E : 0;
        r.lo = valid ? lo : 0;
        r.hi = valid ? hi : 0;
        r.k = k;
        r.flags = (valid ? SEQ_VALID : 0) | flags;
    }
    d2[id] = r;
}

// What the walk needs of one 1024-block: each lane's segment entry and 16
// terms.
struct SeqBlockData {
    SeqDesc sg;
    double v[SEQ_CHUNK];
};

__device__ __forceinline__ void load_block(const double* __restrict__ terms, long npad,
                                           const SeqDesc* __restrict__ dseg, int q, long n1,
                                           long b1, SeqBlockData& b) {
    const int lane = threadIdx.x & 63;
    b.sg = dseg[((long)q * n1 + b1) * 64 + lane];
    load_chunk(terms + (long)q * npad + b1 * SEQ_BLOCK1 + lane * SEQ_CHUNK, b.v);
}

// Terms [a, b) of the block staged in LDS (a, b multiples of 16), added one
// by one: all 16 reads of a chunk are issued before its adds.
__device__ __forceinline__ void seq_chain_lds(const double* blk, int a, int b, long long& sb) {
    double s = __longlong_as_double(sb);
    for (int e = a; e < b; e += SEQ_CHUNK) {
        double2 w[SEQ_CHUNK / 2];
#pragma unroll
        for (int u = 0; u < SEQ_CHUNK / 2; ++u) w[u] = reinterpret_cast<const double2*>(blk + e)[u];
#pragma unroll
        for (int u = 0; u < SEQ_CHUNK / 2; ++u) {
            s = s + w[u].x;
            s = s + w[u].y;
        }
    }
    sb = __double_as_longlong(s);
}

// Stage an opened block's terms in LDS.  Only one wavefront of the
// workgroup walks, and its LDS accesses run in order, so ordering them needs
// no barrier (a barrier would also wait for the load of the next block).
__device__ __forceinline__ void seq_stage(const SeqBlockData& b, double* blk) {
    const int lane = threadIdx.x & 63;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK / 2; ++e)
        reinterpret_cast<double2*>(blk + lane * SEQ_CHUNK)[e] = make_double2(b.v[2 * e], b.v[2 * e + 1]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// All 1024 staged terms added one by one, the reads of the next 16 in
// flight while the current 16 are added.
__device__ __forceinline__ void seq_chain_block(const double* blk, long long& sb) {
    double s = __longlong_as_double(sb);
    const double2* b2 = reinterpret_cast<const double2*>(blk);
    double2 a[SEQ_CHUNK / 2], b[SEQ_CHUNK / 2];
#pragma unroll
    for (int u = 0; u < SEQ_CHUNK / 2; ++u) a[u] = b2[u];
    for (int e = 0; e < SEQ_BLOCK1 / 2; e += SEQ_CHUNK) {
#pragma unroll
        for (int u = 0; u < SEQ_CHUNK / 2; ++u) b[u] = b2[e + SEQ_CHUNK / 2 + u];
#pragma unroll
        for (int u = 0; u < SEQ_CHUNK / 2; ++u) {
            s = s + a[u].x;
            s = s + a[u].y;
        }
        if (e + SEQ_CHUNK < SEQ_BLOCK1 / 2) {
#pragma unroll
            for (int u = 0; u < SEQ_CHUNK / 2; ++u) a[u] = b2[e + SEQ_CHUNK + u];
        }
#pragma unroll
        for (int u = 0; u < SEQ_CHUNK / 2; ++u) {
            s = s + b[u].x;
            s = s + b[u].y;
        }
    }
    sb = __double_as_longlong(s);
}

// Walk a staged block by its segments: a segment of three chunks or more is
// tried as a whole; one that does not apply, and any shorter one, is added
// term by term (a shorter check costs more than the adds it saves).  A
// block cut into more than four segments is added term by term outright:
// its segments are short and their checks would cost more than the adds.
__device__ __forceinline__ void seq_walk_segments(const SeqDesc& sgl, const double* blk,
                                                  long long& sb, long long (&cnt)[9], bool want) {
    if (__popcll(__ballot((sgl.flags & SEQ_START) != 0)) > 4) {
        const long long t_chain = want ? clock64() : 0;
        seq_chain_block(blk, sb);
        cnt[4] += 64;
        if (want) cnt[7] += clock64() - t_chain;
        return;
    }
    for (int c = 0; c < 64;) {
        SeqDesc sg = rldesc(sgl, c);
        const int end = sg.flags >> 8;
        sg.flags &= SEQ_VALID | SEQ_NEG | SEQ_ZERO;
        if (end - c >= 3 && seq_apply(sg, sb)) {
            cnt[2] += 1;
        } else {
            const long long t_chain = want ? clock64() : 0;
            seq_chain_lds(blk, c * SEQ_CHUNK, end * SEQ_CHUNK, sb);
            cnt[4] += end - c;
            if (want) cnt[7] += clock64() - t_chain;
        }
        c = end;
    }
}

// The running sum of sequence blockIdx.x.  A sequence whose 1024-blocks
// mostly leave their binade is added by the plain chain (all four
// wavefronts); otherwise one wavefront walks it.  Every lane holds the same
// s, kept as its bits; descriptors are
// loaded a row at a time and read lane by lane.  A 1024-block whose own
// entry does not apply is staged in LDS and walked by segments, then
// chunks, then single terms.  Inside a 16384-block that does not apply, the
// next 1024-block is loaded into one of two register sets while the block in
// the other is tried and walked.
__global__ void __launch_bounds__(256) seq_walk_kernel(const double* __restrict__ terms, long npad,
                                                       const SeqDesc* __restrict__ dseg,
                                                       const SeqDesc* __restrict__ d1,
                                                       const SeqDesc* __restrict__ d2,
                                                       double* __restrict__ sums,
                                                       long long* __restrict__ stats) {
    __shared__ double buf[2][SEQ_CHAIN_ROW];
    __shared__ int s_bad[4];
    double* blk = buf[0];
    const int q = blockIdx.x;
    // how many of the sequence's 1024-blocks are not of one binade
    {
        const long nb1 = npad / SEQ_BLOCK1;
        int bad = 0;
        for (long b = threadIdx.x; b < nb1; b += 256)
            bad += (d1[(long)q * nb1 + b].flags & (SEQ_VALID | SEQ_ZERO)) == 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o, 64);
        if ((threadIdx.x & 63) == 0) s_bad[threadIdx.x >> 6] = bad;
        __syncthreads();
    }
    const long n1_ = npad / SEQ_BLOCK1;
    if (2L * (s_bad[0] + s_bad[1] + s_bad[2] + s_bad[3]) >= n1_) {
        // at least half the 1024-blocks leave their binade: the walk would
        // add most terms one by one at a slower pace than the plain chain
        const long long t0 = stats ? clock64() : 0;
        const double s = seq_chain_group(terms + (long)q * npad, npad, buf);
        if (threadIdx.x == 0) {
            sums[2 * q] = s;
            sums[2 * q + 1] = 0.0;
            if (stats) {
                for (int u = 0; u < 9; ++u) stats[9 * q + u] = 0;
                stats[9 * q + 4] = npad / SEQ_CHUNK;
                stats[9 * q + 6] = clock64() - t0;
            }
        }
        return;
    }
    if (threadIdx.x >= 64) return;  // the walk is one wavefront's
    const int lane = threadIdx.x;
    // counters (only when stats != nullptr): applied 16384-blocks, 1024-blocks,
    // segments, chunks; chunks added term by term; 1024-blocks opened; clock
    // cycles in total, in term-by-term adds, in opened 1024-blocks
    long long cnt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool want = stats != nullptr;
    const long long t_start = want ? clock64() : 0;
    const long n1 = npad / SEQ_BLOCK1, n2 = npad / SEQ_BLOCK2;
    SeqDesc none;
    none.D = none.lo = none.hi = 0;
    none.k = 0;
    none.flags = 0;
    long long sb = 0;  // +0.0
    SeqBlockData bufA, bufB;
    for (long base2 = 0; base2 < n2; base2 += 64) {
        const SeqDesc my2 = base2 + lane < n2 ? d2[(long)q * n2 + base2 + lane] : none;
        const int cnt2 = n2 - base2 < 64 ? (int)(n2 - base2) : 64;
        for (int i = 0; i < cnt2; ++i) {
            if (seq_apply(rldesc(my2, i), sb)) {
                cnt[0] += 1;
                continue;
            }
            const long base1 = (base2 + i) * 16;
            const SeqDesc my1 = lane < 16 ? d1[(long)q * n1 + base1 + lane] : none;
            // block j's data is in bufA when inA, else in bufB; block j + 1 is
            // loaded into the other set before block j is tried
            load_block(terms, npad, dseg, q, n1, base1, bufA);
            bool inA = true;
            for (int j = 0; j < 16; ++j) {
                if (j + 1 < 16) {
                    if (inA)
                        load_block(terms, npad, dseg, q, n1, base1 + j + 1, bufB);
                    else
                        load_block(terms, npad, dseg, q, n1, base1 + j + 1, bufA);
                }
                if (seq_apply(rldesc(my1, j), sb)) {
                    cnt[1] += 1;
                    inA = !inA;
                    continue;
                }
                const long long t_open = want ? clock64() : 0;
                cnt[5] += 1;
                SeqDesc sgl;
                if (inA) {
                    seq_stage(bufA, blk);
                    sgl = bufA.sg;
                } else {
                    seq_stage(bufB, blk);
                    sgl = bufB.sg;
                }
                inA = !inA;
                seq_walk_segments(sgl, blk, sb, cnt, want);
                if (want) cnt[8] += clock64() - t_open;
            }
        }
    }
    if (lane == 0) {
        sums[2 * q] = __longlong_as_double(sb);
        sums[2 * q + 1] = 0.0;
    }
    if (want && lane == 0) {
        cnt[6] = clock64() - t_start;
        for (int u = 0; u < 9; ++u) stats[9 * q + u] = cnt[u];
    }
}

}  // namespace

long seqsum_npad(long n) { return (n + SEQ_BLOCK2 - 1) / SEQ_BLOCK2 * SEQ_BLOCK2; }

long seqsum_workspace_len(long npad, int nq) {
    const long n0 = npad / SEQ_CHUNK, n1 = npad / SEQ_BLOCK1, n2 = npad / SEQ_BLOCK2;
    const long dsz = (long)(sizeof(SeqDesc) / sizeof(double));
    return (long)nq * (2 * n1 + dsz * (n0 + n1 + n2));
}

hipError_t launch_seqsum(int method, const double* terms, long npad, int nq, double* ws,
                         double* sums, hipStream_t st, long long* stats) {
    if (npad <= 0 || npad % SEQ_BLOCK2 != 0 || nq <= 0) return hipErrorInvalidValue;
    if (method == SEQSUM_CHAIN) {
        hipLaunchKernelGGL(seq_chain_kernel, dim3((unsigned)nq), dim3(256), 0, st, terms, npad,
                           sums);
        return hipGetLastError();
    }
    if (method != SEQSUM_COMPOSITE) return hipErrorInvalidValue;
    const long n0 = npad / SEQ_CHUNK, n1 = npad / SEQ_BLOCK1, n2 = npad / SEQ_BLOCK2;
    double* bs = ws;
    double* pre = bs + (long)nq * n1;
    SeqDesc* dseg = reinterpret_cast<SeqDesc*>(pre + (long)nq * n1);
    SeqDesc* d1 = dseg + (long)nq * n0;
    SeqDesc* d2 = d1 + (long)nq * n1;
    const dim3 g1((unsigned)((n1 + 3) / 4), (unsigned)nq);
    hipLaunchKernelGGL(seq_blocksum_kernel, g1, dim3(256), 0, st, terms, npad, bs);
    hipLaunchKernelGGL(seq_scan_kernel, dim3((unsigned)nq), dim3(1024), 0, st, bs, n1, pre);
    hipLaunchKernelGGL(seq_prep_kernel, g1, dim3(256), 0, st, terms, npad, pre, dseg, d1);
    hipLaunchKernelGGL(seq_level2_kernel, dim3((unsigned)((n2 * nq + 255) / 256)), dim3(256), 0,
                       st, d1, n2, nq, d2);
    hipLaunchKernelGGL(seq_walk_kernel, dim3((unsigned)nq), dim3(256), 0, st, terms, npad, dseg, d1,
                       d2, sums, stats);
    return hipGetLastError();
}

}  // namespace hyd
