// seqsum.hip -- sequential-order sums (see seqsum.hpp for the method).
//
// SEQSUM_PEAKS, per sequence q of npad terms (nb = npad / 1024 blocks,
// nc = npad / 4096 chunks):
//   pk_blocksum_kernel  double-double sum of every 1024-term block
//   pk_scan_kernel      exclusive prefix of the block sums (double-double)
//   pk_peak_kernel      per 4096-term block: the predicted running sum after
//                       each term, where its magnitude peaks (chunk starts),
//                       and whether every term is zero
//   pk_tab_kernel       one start value per residue of each chunk, run over
//                       the chunk's terms: 2 lanes per chunk for P = 1, 8
//                       for P = 2, 3, a wavefront for P = 4..6
//   pk_stitch_kernel    one wavefront per sequence carries the exact sum
//                       across the chunks (term by term where a chunk fails)
//   pk_verify_kernel    one lane per chunk runs it again from its exact start
//   pk_redo_kernel      the chain, only for a sequence whose check failed
#include "seqsum.hpp"

namespace hyd {

namespace {

constexpr int SEQ_CHAIN_TILE = 2048;
constexpr int SEQ_CHAIN_ROW = SEQ_CHAIN_TILE + 32;  // a tile and the overread of its last reads
constexpr int TAB = 1 << SEQ_PMAX;                  // tabulated end values per chunk
constexpr int BADP = 99;                            // a chunk that is added term by term

__device__ __forceinline__ unsigned long long ubits(double d) {
    return (unsigned long long)__double_as_longlong(d);
}
__device__ __forceinline__ double dfrom(unsigned long long u) {
    return __longlong_as_double((long long)u);
}
// sign and exponent field: equal keys = same sign and binade
__device__ __forceinline__ unsigned key_of(double d) { return (unsigned)(ubits(d) >> 52); }
__device__ __forceinline__ int expo(double d) { return (int)((ubits(d) >> 52) & 0x7ff); }
__device__ __forceinline__ bool normal_finite(double d) {
    const int e = expo(d);
    return e != 0 && e != 0x7ff;
}
// s = I ulp(s): the signed integer I (2^52 <= |I| < 2^53) of a normal double
__device__ __forceinline__ long long imant(double d) {
    const unsigned long long u = ubits(d);
    const long long m = (long long)((u & ((1ull << 52) - 1)) | (1ull << 52));
    return (u >> 63) ? -m : m;
}
__device__ __forceinline__ double with_imant(long long i, double like) {
    const unsigned long long a = (unsigned long long)(i < 0 ? -i : i);
    return dfrom((ubits(like) & (0x7ffull << 52)) | (a & ((1ull << 52) - 1)) |
                 (i < 0 ? (1ull << 63) : 0ull));
}
// The start value of residue r (I mod 2^P) nearest below V's integer, kept
// inside V's binade.
__device__ __forceinline__ double variant_start(double V, int P, int r) {
    const long long mod = 1ll << P;
    long long i = imant(V);
    i = i - (i & (mod - 1)) + r;
    const long long a = i < 0 ? -i : i;
    if (a < (1ll << 52)) i += i > 0 ? mod : -mod;
    else if (a >= (1ll << 53)) i -= i > 0 ? mod : -mod;
    return with_imant(i, V);
}
// Residue bits of a chunk from its predicted start (peak) V and the next
// chunk's predicted peak V1: 1 + how far V1's binade lies above V's, or BADP.
__device__ __forceinline__ int chunk_bits(double V, double V1) {
    if (!normal_finite(V) || !normal_finite(V1)) return BADP;
    const int d = expo(V1) - expo(V);
    const int P = (d > 0 ? d : 0) + 1;
    return P <= SEQ_PMAX ? P : BADP;
}

// Reads of one lane's value into scalar registers (the lane index uniform).
__device__ __forceinline__ double rl_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ long rl_l(long v, int lane) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffl), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return (long)(((unsigned long)hi << 32) | lo);
}
__device__ __forceinline__ unsigned long long rfl_u(unsigned long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double rfl_d(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b) {
    DD r;
    r.hi = a + b;
    const double bb = r.hi - a;
    r.lo = (a - (r.hi - bb)) + (b - bb);
    return r;
}
// double-double addition for predictions (not error-free in the last bits)
__device__ __forceinline__ DD dd_add(DD x, DD y) {
    const DD s = two_sum(x.hi, y.hi);
    return two_sum(s.hi, s.lo + (x.lo + y.lo));
}
__device__ __forceinline__ DD dd_shfl_xor(DD x, int o) {
    return DD{__shfl_xor(x.hi, o, 64), __shfl_xor(x.lo, o, 64)};
}
__device__ __forceinline__ DD dd_shfl_up(DD x, int o) {
    return DD{__shfl_up(x.hi, o, 64), __shfl_up(x.lo, o, 64)};
}

struct Peak {
    double V;  // predicted running sum at the peak (start of the chunk)
    int p;     // position of the peak's term in its 4096-term block
    int zero;  // 1 when every term of the block is +0 or -0
};

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

// Double-double sum of each 1024-term block (a wavefront per block).
__global__ void __launch_bounds__(256) pk_blocksum_kernel(const double* __restrict__ terms,
                                                          long npad, double* __restrict__ bsum) {
    const long nb = npad / SEQ_BLOCK1;
    const long j = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= nb) return;
    const int q = blockIdx.y;
    double v[SEQ_CHUNK];
    load_chunk(terms + (long)q * npad + j * SEQ_BLOCK1 + lane * SEQ_CHUNK, v);
    DD acc{0.0, 0.0};
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) acc = dd_add(acc, DD{v[e], 0.0});
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc = dd_add(acc, dd_shfl_xor(acc, o));
    if (lane == 0) {
        bsum[2 * ((long)q * nb + j)] = acc.hi;
        bsum[2 * ((long)q * nb + j) + 1] = acc.lo;
    }
}

// Exclusive prefix of the nb block sums of sequence blockIdx.x.
__global__ void __launch_bounds__(1024) pk_scan_kernel(const double* __restrict__ bsum, long nb,
                                                       double* __restrict__ bpre) {
    __shared__ double sh[1024], sl[1024];
    const int tid = threadIdx.x;
    const long per = (nb + 1023) / 1024;
    const long b = (long)tid * per;
    const long e = b + per < nb ? b + per : nb;
    const double* in = bsum + 2 * (long)blockIdx.x * nb;
    double* out = bpre + 2 * (long)blockIdx.x * nb;
    DD t{0.0, 0.0};
    for (long i = b; i < e; ++i) t = dd_add(t, DD{in[2 * i], in[2 * i + 1]});
    sh[tid] = t.hi;
    sl[tid] = t.lo;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        DD add{0.0, 0.0};
        if (tid >= o) add = DD{sh[tid - o], sl[tid - o]};
        __syncthreads();
        if (tid >= o) {
            const DD c = dd_add(DD{sh[tid], sl[tid]}, add);
            sh[tid] = c.hi;
            sl[tid] = c.lo;
        }
        __syncthreads();
    }
    DD run{0.0, 0.0};
    if (tid > 0) run = DD{sh[tid - 1], sl[tid - 1]};
    for (long i = b; i < e; ++i) {
        out[2 * i] = run.hi;
        out[2 * i + 1] = run.lo;
        run = dd_add(run, DD{in[2 * i], in[2 * i + 1]});
    }
}

// Per 4096-term block (a workgroup, a wavefront per 1024 terms): the
// predicted running sum after each term (block prefix plus the in-block
// prefix), the first position of its largest magnitude, and whether all the
// terms are zero.
__global__ void __launch_bounds__(256) pk_peak_kernel(const double* __restrict__ terms, long npad,
                                                      const double* __restrict__ bpre,
                                                      Peak* __restrict__ pk) {
    __shared__ double s_best[4], s_v[4];
    __shared__ int s_k[4], s_zero[4];
    const long nb = npad / SEQ_BLOCK1;
    const int w = threadIdx.x >> 6;
    const long j = (long)blockIdx.x * 4 + w;  // < nb: nb is a multiple of 16
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.y;
    double v[SEQ_CHUNK];
    load_chunk(terms + (long)q * npad + j * SEQ_BLOCK1 + lane * SEQ_CHUNK, v);
    double pre[SEQ_CHUNK];
    double run = 0.0;
    bool zero = true;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        run = run + v[e];
        pre[e] = run;
        zero = zero && v[e] == 0.0;
    }
    DD incl{run, 0.0};
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const DD u = dd_shfl_up(incl, o);
        if (lane >= o) incl = dd_add(incl, u);
    }
    DD excl = dd_shfl_up(incl, 1);
    if (lane == 0) excl = DD{0.0, 0.0};
    const long bi = (long)q * nb + j;
    const DD base = dd_add(DD{bpre[2 * bi], bpre[2 * bi + 1]}, excl);
    double best = -1.0, bv = 0.0;
    int bk = 0;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        const double pv = base.hi + (base.lo + pre[e]);
        const double a = fabs(pv);
        if (a > best) {
            best = a;
            bv = pv;
            bk = w * SEQ_BLOCK1 + lane * SEQ_CHUNK + e;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64), ov = __shfl_xor(bv, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ob > best || (ob == best && ok < bk)) {
            best = ob;
            bv = ov;
            bk = ok;
        }
    }
    const int wzero = __all(zero);
    if (lane == 0) {
        s_best[w] = best;
        s_v[w] = bv;
        s_k[w] = bk;
        s_zero[w] = wzero;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // in term order: a later wavefront wins only with a larger magnitude
        Peak r;
        r.V = s_v[0];
        r.p = s_k[0];
        r.zero = s_zero[0];
        double b = s_best[0];
        for (int u = 1; u < 4; ++u) {
            if (s_best[u] > b) {
                b = s_best[u];
                r.V = s_v[u];
                r.p = s_k[u];
            }
            r.zero = r.zero && s_zero[u];
        }
        pk[(long)q * gridDim.x + blockIdx.x] = r;
    }
}

// Chunk c of a sequence: terms [first, last], residue bits P, predicted
// start V, and whether all its terms are zero (both 4096-blocks it touches).
struct Chunk {
    long first, last;
    double V;
    int P;
    int zero;
};
__device__ __forceinline__ Chunk chunk_of(const Peak* __restrict__ pkq, long nc, long npad,
                                          long c) {
    const Peak a = pkq[c];
    Chunk k;
    k.first = c * SEQ_CBLK + a.p + 1;
    k.V = a.V;
    if (c + 1 < nc) {
        const Peak b = pkq[c + 1];
        k.last = (c + 1) * SEQ_CBLK + b.p;
        k.P = chunk_bits(a.V, b.V);
        k.zero = a.zero && b.zero;
    } else {
        k.last = npad - 1;
        k.P = chunk_bits(a.V, a.V);
        k.zero = a.zero;
    }
    return k;
}

__device__ __forceinline__ long wave_max_long(long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// s + x[first] + ... + x[first + len - 1], one lane's own run, the next 8
// terms loaded while the current 8 are added.  Every lane of the wavefront
// calls it: minlen and maxlen are the shortest and longest runs among the
// lanes (a lane with nothing to add passes len = maxlen and any start).
// Reads reach x[first + maxlen + 15]: the term arrays are followed by at
// least that much readable memory (the workspace).
__device__ __forceinline__ double run_lane(const double* __restrict__ x, long first, long len,
                                           long minlen, long maxlen, double s) {
    const double* p = x + first;
    double cur[8], nxt[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cur[i] = p[i];
    long base = 0;
    for (; base + 8 <= minlen; base += 8) {  // every lane adds all 8
#pragma unroll
        for (int i = 0; i < 8; ++i) nxt[i] = p[base + 8 + i];
#pragma unroll
        for (int i = 0; i < 8; ++i) s = s + cur[i];
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[i] = nxt[i];
    }
    for (; base < maxlen; base += 8) {  // the tails, predicated per lane
#pragma unroll
        for (int i = 0; i < 8; ++i) nxt[i] = p[base + 8 + i];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double t = s + cur[i];
            s = (base + i < len) ? t : s;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[i] = nxt[i];
    }
    return s;
}

__device__ __forceinline__ long wave_min_long(long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// End values of the chunks whose residue bits P lie in [PLO, PHI]: L lanes
// per chunk (lane % L = residue), 64 / L chunks per wavefront.
template <int L, int PLO, int PHI>
__global__ void __launch_bounds__(256) pk_tab_kernel(const double* __restrict__ terms, long npad,
                                                     const Peak* __restrict__ pk,
                                                     double* __restrict__ tab) {
    constexpr int CPW = 64 / L;
    const long nc = npad / SEQ_CBLK;
    const int lane = threadIdx.x & 63;
    const long c = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * CPW + lane / L;
    const int r = lane % L;
    const int q = blockIdx.y;
    bool active = false;
    Chunk k{0, 0, 0.0, BADP, 0};
    if (c < nc) {
        k = chunk_of(pk + (long)q * nc, nc, npad, c);
        active = k.P >= PLO && k.P <= PHI && r < (1 << k.P);
    }
    const long len = active ? k.last - k.first + 1 : 0;
    const long maxlen = wave_max_long(len);
    if (maxlen == 0) return;  // uniform across the wavefront
    const long minlen = wave_min_long(active ? len : maxlen);
    double s = active ? variant_start(k.V, k.P, r) : 0.0;
    s = run_lane(terms + (long)q * npad, active ? k.first : 0, active ? len : maxlen, minlen,
                 maxlen, s);
    if (active) tab[((long)q * nc + c) * TAB + r] = s;
}

// s + x[first] + ... in order: the wavefront loads 64 terms at a time (the
// next 64 while the current ones are added) and the chain of additions runs
// on the values read from lane i.  Returns the same value in every lane.
__device__ __noinline__ double run_wave(const double* __restrict__ x, long first, long len,
                                        double s) {
    const int lane = threadIdx.x & 63;
    double v = lane < len ? x[first + lane] : 0.0;
    for (long base = 0; base < len; base += 64) {
        const long cnt = len - base < 64 ? len - base : 64;
        const double w = base + 64 + lane < len ? x[first + base + 64 + lane] : 0.0;
        if (cnt == 64) {
#pragma unroll
            for (int i = 0; i < 64; ++i) s = s + rl_d(v, i);
        } else {
            for (int i = 0; i < (int)cnt; ++i) s = s + rl_d(v, i);
        }
        v = w;
    }
    return rfl_d(s);
}

constexpr int STITCH_G = 16;        // chunks per load round of the stitch
// LDS bytes asked for each stitch workgroup (dynamic, unused): more than half
// of a CU's 160 KB, so no two stitches share a CU
constexpr int STITCH_LDS_BYTES = 96 * 1024;

// What the stitch needs of a chunk: its start key (sign and exponent field),
// residue bits, the start values' common base I_v - (I_v mod 2^P), and
// whether base + r stays in the binade for every residue r.
struct StitchInfo {
    unsigned key;
    int P;
    long long base;
    int ok;
    int zero;
};
__device__ __forceinline__ StitchInfo stitch_info(const Chunk& k) {
    StitchInfo i;
    i.key = key_of(k.V);
    i.P = k.P;
    i.zero = k.zero;
    i.ok = k.P <= SEQ_PMAX;
    i.base = 0;
    if (i.ok) {
        const long long mod = 1ll << k.P;
        const long long iv = imant(k.V);
        i.base = iv - (iv & (mod - 1));
        const long long lo = i.base, hi = i.base + (mod - 1);
        const long long alo = lo < 0 ? -lo : lo, ahi = hi < 0 ? -hi : hi;
        i.ok = alo >= (1ll << 52) && alo < (1ll << 53) && ahi >= (1ll << 52) && ahi < (1ll << 53);
    }
    return i;
}

// One chunk of the stitch in integer arithmetic on the bits u of the exact
// start: r = I mod 2^P, delta = I - (base + r) (a multiple of 2^P), and the
// end E[r] + delta ulp(start) as the integer of E[r]'s binade.  False when
// the start is not in the predicted binade or the end leaves E[r]'s.
__device__ __forceinline__ bool stitch_step(unsigned long long& u, unsigned key, int P,
                                            long long base, double Eg) {
    if ((unsigned)(u >> 52) != key) return false;
    const long long m = (long long)((u & ((1ull << 52) - 1)) | (1ull << 52));
    const long long I = (u >> 63) ? -m : m;
    const int r = (int)(I & ((1ll << P) - 1));
    const long long delta = I - base - r;
    const unsigned long long e = ubits(rl_d(Eg, r));
    const int ee = (int)((e >> 52) & 0x7ff), es = (int)((u >> 52) & 0x7ff);
    if (ee == 0 || ee == 0x7ff) return false;
    const long long me = (long long)((e & ((1ull << 52) - 1)) | (1ull << 52));
    const long long IE = (e >> 63) ? -me : me;
    const int k = es - ee;
    long long D;
    if (k >= 0) {
        if (k > 10 || delta >= (1ll << 50) || delta <= -(1ll << 50)) return false;
        D = delta * (1ll << k);
    } else {
        D = delta >> (-k);  // exact: delta is a multiple of 2^P and -k < P
    }
    const long long Ie = IE + D;
    if ((Ie < 0) != (IE < 0)) return false;
    const long long a = Ie < 0 ? -Ie : Ie;
    if (a < (1ll << 52) || a >= (1ll << 53)) return false;
    u = rfl_u((e & (0xfffull << 52)) | ((unsigned long long)a & ((1ull << 52) - 1)));
    return true;
}

// One wavefront per sequence carries the exact running sum from chunk to
// chunk.  Lane r holds tab[c][r] of the STITCH_G chunks in hand, so E[r] is a
// read of lane r; the next round's tables load meanwhile.  A chunk of zeros
// leaves the sum as it is: s + 0 and s - 0 are s for every s but -0 and NaN,
// and the sum, which starts at +0, is never -0 (the last bits are checked by
// the verification anyway).  The launch asks
// for STITCH_LDS_BYTES of LDS so that each stitch has a CU of its own (they
// are serial chains of scalar work that would otherwise share a CU's scalar
// unit).
__global__ void __launch_bounds__(64) pk_stitch_kernel(const double* __restrict__ terms, long npad,
                                                       const Peak* __restrict__ pk,
                                                       const double* __restrict__ tab,
                                                       double* __restrict__ starts,
                                                       unsigned long long* __restrict__ flags,
                                                       double* __restrict__ sums,
                                                       long long* __restrict__ stats) {
    const long nc = npad / SEQ_CBLK;
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const double* x = terms + (long)q * npad;
    const Peak* pkq = pk + (long)q * nc;
    const double* tq = tab + (long)q * nc * TAB;
    double* sq = starts + (long)q * (nc + 1);
    if (lane == 0) flags[q] = 0ull;
    const bool want = stats != nullptr;
    const long long t0 = want ? (long long)clock64() : 0;
    long long n_small = 0, n_big = 0, n_fall = 0, cyc_fall = 0, n_zero = 0;
    // the head: the terms up to the first block's peak, from 0
    // (run_wave returns the same value in every lane; readfirstlane tells the
    // compiler so, which keeps the stitch's integer work on the scalar unit)
    unsigned long long u = rfl_u(ubits(run_wave(x, 0, pkq[0].p + 1, 0.0)));
    double E[STITCH_G], En[STITCH_G];
#pragma unroll
    for (int g = 0; g < STITCH_G; ++g) E[g] = g < nc ? tq[(long)g * TAB + lane] : 0.0;
    Chunk mine = chunk_of(pkq, nc, npad, lane < nc ? lane : 0);
    for (long c0 = 0; c0 < nc; c0 += STITCH_G) {
        const int g_n = nc - c0 < STITCH_G ? (int)(nc - c0) : STITCH_G;
        const long c1 = c0 + STITCH_G;
        // the next round's tables and chunk bounds, in flight during this round
#pragma unroll
        for (int g = 0; g < STITCH_G; ++g) En[g] = c1 + g < nc ? tq[(c1 + g) * TAB + lane] : 0.0;
        const Chunk next = chunk_of(pkq, nc, npad, c1 + lane < nc ? c1 + lane : 0);
        const StitchInfo info = stitch_info(mine);
        double mystart = 0.0;
#pragma unroll
        for (int g = 0; g < STITCH_G; ++g) {
            if (g < g_n) {
                if (lane == g) mystart = __longlong_as_double((long long)u);
                const int P = __builtin_amdgcn_readlane(info.P, g);
                const bool ok = __builtin_amdgcn_readlane(info.ok, g) != 0;
                const unsigned key = (unsigned)__builtin_amdgcn_readlane((int)info.key, g);
                const long long base = rl_l(info.base, g);
                const bool zero = __builtin_amdgcn_readlane(info.zero, g) != 0;
                const bool nan = (u & (0x7ffull << 52)) == (0x7ffull << 52) &&
                                 (u & ((1ull << 52) - 1)) != 0;
                if (P <= 3) ++n_small;
                else if (P <= SEQ_PMAX) ++n_big;
                if (zero && !nan) {
                    ++n_zero;
                } else if (!(ok && stitch_step(u, key, P, base, E[g]))) {
                    const long first = rl_l(mine.first, g);
                    const long last = rl_l(mine.last, g);
                    const long long tf = want ? (long long)clock64() : 0;
                    u = rfl_u(ubits(run_wave(x, first, last - first + 1,
                                             __longlong_as_double((long long)u))));
                    if (want) cyc_fall += (long long)clock64() - tf;
                    ++n_fall;
                }
            }
        }
        if (lane < g_n) sq[c0 + lane] = mystart;
#pragma unroll
        for (int g = 0; g < STITCH_G; ++g) E[g] = En[g];
        mine = next;
    }
    const double s = __longlong_as_double((long long)u);
    if (lane == 0) {
        sq[nc] = s;
        sums[2 * q] = s;
        sums[2 * q + 1] = 0.0;
        if (want) {
            long long* o = stats + (long)SEQ_STATS * q;
            o[0] = nc;
            o[1] = n_small;
            o[2] = n_big;
            o[3] = n_fall;
            o[4] = 0;
            o[5] = (long long)clock64() - t0;
            o[6] = cyc_fall;
            o[7] = 0;
            o[8] = n_zero;
        }
    }
}

// One lane per chunk: run it from its exact start; it must reach the next
// chunk's start.  A miss marks the sequence for the chain.
__global__ void __launch_bounds__(256) pk_verify_kernel(const double* __restrict__ terms,
                                                        long npad, const Peak* __restrict__ pk,
                                                        const double* __restrict__ starts,
                                                        unsigned long long* __restrict__ flags,
                                                        long long* __restrict__ stats) {
    const long nc = npad / SEQ_CBLK;
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const int q = blockIdx.y;
    const double* sq = starts + (long)q * (nc + 1);
    Chunk k{0, 0, 0.0, 0, 0};
    long len = 0;
    if (c < nc) {
        k = chunk_of(pk + (long)q * nc, nc, npad, c);
        len = k.last - k.first + 1;
    }
    const long maxlen = wave_max_long(len);
    if (maxlen == 0) return;  // uniform across the wavefront
    const long minlen = wave_min_long(c < nc ? len : maxlen);
    const double s = run_lane(terms + (long)q * npad, c < nc ? k.first : 0,
                              c < nc ? len : maxlen, minlen, maxlen, c < nc ? sq[c] : 0.0);
    if (c < nc && ubits(s) != ubits(sq[c + 1])) {
        flags[q] = 1ull;
        if (stats) atomicAdd((unsigned long long*)&stats[(long)SEQ_STATS * q + 4], 1ull);
    }
}

// The chain, for the sequences whose verification failed.
__global__ void __launch_bounds__(256) pk_redo_kernel(const double* __restrict__ terms, long npad,
                                                      const unsigned long long* __restrict__ flags,
                                                      double* __restrict__ sums,
                                                      long long* __restrict__ stats) {
    __shared__ double buf[2][SEQ_CHAIN_ROW];
    const int q = blockIdx.x;
    if (flags[q] == 0ull) return;  // uniform across the workgroup
    const double s = seq_chain_group(terms + (long)q * npad, npad, buf);
    if (threadIdx.x == 0) {
        sums[2 * q] = s;
        sums[2 * q + 1] = 0.0;
        if (stats) stats[(long)SEQ_STATS * q + 7] = 1;
    }
}

}  // namespace

long seqsum_npad(long n) { return (n + SEQ_BLOCK2 - 1) / SEQ_BLOCK2 * SEQ_BLOCK2; }

// bsum and bpre (2 nb each), peaks (2 nc), tables (64 nc), chunk starts
// (nc + 1) and a flag per sequence
long seqsum_workspace_len(long npad, int nq) {
    const long nb = npad / SEQ_BLOCK1, nc = npad / SEQ_CBLK;
    return (long)nq * (4 * nb + 2 * nc + (long)TAB * nc + nc + 2);
}

hipError_t launch_seqsum(int method, const double* terms, long npad, int nq, double* ws,
                         double* sums, hipStream_t st, long long* stats) {
    if (npad <= 0 || npad % SEQ_BLOCK2 != 0 || nq <= 0) return hipErrorInvalidValue;
    if (method == SEQSUM_CHAIN) {
        hipLaunchKernelGGL(seq_chain_kernel, dim3((unsigned)nq), dim3(256), 0, st, terms, npad,
                           sums);
        return hipGetLastError();
    }
    if (method != SEQSUM_PEAKS) return hipErrorInvalidValue;
    // the chunk runs read up to 16 doubles past a sequence: the workspace
    // must follow the terms
    if (ws < terms + (long)nq * npad) return hipErrorInvalidValue;
    const long nb = npad / SEQ_BLOCK1, nc = npad / SEQ_CBLK;
    double* bsum = ws;
    double* bpre = bsum + 2 * (long)nq * nb;
    Peak* pk = reinterpret_cast<Peak*>(bpre + 2 * (long)nq * nb);
    double* tab = reinterpret_cast<double*>(pk + (long)nq * nc);
    double* starts = tab + (long)TAB * nq * nc;
    unsigned long long* flags =
        reinterpret_cast<unsigned long long*>(starts + (long)nq * (nc + 1));
    const dim3 gb((unsigned)(nb / 4), (unsigned)nq);  // 4 wavefronts per 4096-block
    hipLaunchKernelGGL(pk_blocksum_kernel, gb, dim3(256), 0, st, terms, npad, bsum);
    hipLaunchKernelGGL(pk_scan_kernel, dim3((unsigned)nq), dim3(1024), 0, st, bsum, nb, bpre);
    hipLaunchKernelGGL(pk_peak_kernel, gb, dim3(256), 0, st, terms, npad, bpre, pk);
    hipLaunchKernelGGL((pk_tab_kernel<2, 1, 1>), dim3((unsigned)((nc + 127) / 128), (unsigned)nq),
                       dim3(256), 0, st, terms, npad, pk, tab);
    hipLaunchKernelGGL((pk_tab_kernel<8, 2, 3>), dim3((unsigned)((nc + 31) / 32), (unsigned)nq),
                       dim3(256), 0, st, terms, npad, pk, tab);
    hipLaunchKernelGGL((pk_tab_kernel<64, 4, SEQ_PMAX>), dim3((unsigned)((nc + 3) / 4), (unsigned)nq),
                       dim3(256), 0, st, terms, npad, pk, tab);
    hipLaunchKernelGGL(pk_stitch_kernel, dim3((unsigned)nq), dim3(64), STITCH_LDS_BYTES, st, terms,
                       npad, pk, tab, starts, flags, sums, stats);
    hipLaunchKernelGGL(pk_verify_kernel, dim3((unsigned)((nc + 255) / 256), (unsigned)nq),
                       dim3(256), 0, st, terms, npad, pk, starts, flags, stats);
    hipLaunchKernelGGL(pk_redo_kernel, dim3((unsigned)nq), dim3(256), 0, st, terms, npad, flags,
                       sums, stats);
    return hipGetLastError();
}

}  // namespace hyd
