// seqsum.hip -- sequential-order sums (see seqsum.hpp for the method).
//
// SEQSUM_PEAKS, per sequence q of npad terms (nb = npad / 1024 blocks,
// nc = npad / 4096 chunks, ng = nc / 16 groups of chunks):
//   pk_block_kernel     per 1024-term block: double-double sum, largest and
//                       smallest in-block prefix, all-zero flag
//   pk_scan_kernel      prefixes of the block sums and, per 4096-term block,
//                       the predicted peak of the running sum's magnitude
//                       (chunk starts)
//   pk_class_kernel     the list of start values to tabulate, 2^P per chunk,
//                       longest chunks first
//   pk_tab_kernel       a lane per start value runs it through the chunk's
//                       terms: the end value and its exactness certificate
//   pk_group_kernel     per group of 16 chunks and start residue: where the
//                       group ends and how far a start may move (the map)
//   pk_stitch_kernel    one wavefront per sequence carries the exact sum
//                       from group to group through the maps (chunk by
//                       chunk, term by term where checks fail)
// SEQSUM_PEAKS_CHECKED (and any call that asks for counters) also runs
//   pk_fill_kernel      per group: the exact start of every chunk
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

// Per 1024-term block (a wavefront): the double-double sum, the largest and
// smallest in-block prefix (the running sum relative to the block's start)
// with their first positions, and whether every term is zero.  The peak of
// |base + prefix| over a block lies at one of the two extremes, so the
// prefixes of the block sums then place every chunk start without reading
// the terms again.
struct BlockStat {
    double hi, lo;      // block sum
    double vmax, vmin;  // largest and smallest in-block prefix
    int pmax, pmin;     // their first positions
    int zero, pad;
};

__global__ void __launch_bounds__(256) pk_block_kernel(const double* __restrict__ terms, long npad,
                                                       BlockStat* __restrict__ bs) {
    const long nb = npad / SEQ_BLOCK1;
    const long j = (long)blockIdx.x * 4 + (threadIdx.x >> 6);  // < nb: nb is a multiple of 16
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.y;
    double v[SEQ_CHUNK];
    load_chunk(terms + (long)q * npad + j * SEQ_BLOCK1 + lane * SEQ_CHUNK, v);
    // the lane's running sum (its prefixes) with the exact error of each
    // addition accumulated beside it: a compensated sum (Ogita, Rump and
    // Oishi's Sum2), as accurate as a double-double for 16 terms and a third
    // of its cost; the statistics only predict the running sum, whose exact
    // value the certificates establish later
    double pre[SEQ_CHUNK];
    double run = 0.0, err = 0.0;
    bool zero = true;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        const DD t = two_sum(run, v[e]);
        run = t.hi;
        err = err + t.lo;
        pre[e] = run;
        zero = zero && v[e] == 0.0;
    }
    DD incl = two_sum(run, err);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const DD u = dd_shfl_up(incl, o);
        if (lane >= o) incl = dd_add(incl, u);
    }
    DD excl = dd_shfl_up(incl, 1);
    if (lane == 0) excl = DD{0.0, 0.0};
    double vmax = -__builtin_huge_val(), vmin = __builtin_huge_val();
    int pmax = 0, pmin = 0;
#pragma unroll
    for (int e = 0; e < SEQ_CHUNK; ++e) {
        const double rel = excl.hi + (excl.lo + pre[e]);
        if (rel > vmax) {
            vmax = rel;
            pmax = lane * SEQ_CHUNK + e;
        }
        if (rel < vmin) {
            vmin = rel;
            pmin = lane * SEQ_CHUNK + e;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_xor(vmax, o, 64), b = __shfl_xor(vmin, o, 64);
        const int pa = __shfl_xor(pmax, o, 64), pb = __shfl_xor(pmin, o, 64);
        if (a > vmax || (a == vmax && pa < pmax)) {
            vmax = a;
            pmax = pa;
        }
        if (b < vmin || (b == vmin && pb < pmin)) {
            vmin = b;
            pmin = pb;
        }
    }
    const int wz = __all(zero);
    const double thi = __shfl(incl.hi, 63, 64), tlo = __shfl(incl.lo, 63, 64);
    if (lane == 0) {
        BlockStat r;
        r.hi = thi;
        r.lo = tlo;
        r.vmax = vmax;
        r.vmin = vmin;
        r.pmax = pmax;
        r.pmin = pmin;
        r.zero = wz;
        r.pad = 0;
        bs[(long)q * nb + j] = r;
    }
}

// Per sequence (blockIdx.x): the exclusive prefix of the block sums
// (double-double), and per 4096-term block the predicted peak of the running
// sum's magnitude (the first on ties) and whether all its terms are zero.
__global__ void __launch_bounds__(1024) pk_scan_kernel(const BlockStat* __restrict__ bs, long nb,
                                                       Peak* __restrict__ pk) {
    __shared__ double sh[1024], sl[1024];
    const int tid = threadIdx.x;
    // whole 4096-term blocks (four 1024-term blocks) per thread
    const long per = ((nb + 1023) / 1024 + 3) / 4 * 4;
    const long b = (long)tid * per;
    const long e = b + per < nb ? b + per : nb;
    const BlockStat* in = bs + (long)blockIdx.x * nb;
    Peak* out = pk + (long)blockIdx.x * (nb / 4);
    DD t{0.0, 0.0};
    for (long i = b; i < e; ++i) t = dd_add(t, DD{in[i].hi, in[i].lo});
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
    double best = -1.0, bv = 0.0;
    int bp = 0, zero = 1;
    for (long i = b; i < e; ++i) {
        const BlockStat s = in[i];
        const int w = (int)(i & 3);
        if (w == 0) {
            best = -1.0;
            bv = 0.0;
            bp = 0;
            zero = 1;
        }
        const double cmax = run.hi + (run.lo + s.vmax), cmin = run.hi + (run.lo + s.vmin);
        const double amax = fabs(cmax), amin = fabs(cmin);
        const bool lo = amin > amax || (amin == amax && s.pmin < s.pmax);
        const double a = lo ? amin : amax;
        if (a > best) {
            best = a;
            bv = lo ? cmin : cmax;
            bp = w * SEQ_BLOCK1 + (lo ? s.pmin : s.pmax);
        }
        zero = zero && s.zero;
        if (w == 3) {
            Peak r;
            r.V = bv;
            r.p = bp;
            r.zero = zero;
            out[i >> 2] = r;
        }
        run = dd_add(run, DD{s.hi, s.lo});
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
__device__ __forceinline__ long wave_min_long(long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

constexpr int STITCH_G = 16;  // chunks per load round of a chunk-by-chunk stitch
constexpr int GROUP = 16;     // chunks per group of the two-level stitch
constexpr int MAP_BATCH = 8;  // groups whose maps pass through LDS together
constexpr int PG = 6;         // residue bits of a group's start (one per lane)
constexpr int NRES = 1 << PG; // maps per group
// LDS bytes asked for each sequence-level stitch workgroup (the map ring,
// 64 KB, and more): more than half of a CU's 160 KB, so no two share a CU
constexpr int STITCH_LDS_BYTES = 96 * 1024;
constexpr unsigned long long MANT = (1ull << 52) - 1;
constexpr double ALLOW_MAX = 288230376151711744.0;  // 2^58, above any |delta|

// What a stitch needs of a chunk: its start key (sign and exponent field),
// residue bits, the start values' common base I_v - (I_v mod 2^P), whether
// base + r stays in the binade for every residue r, and whether all its
// terms are zero.
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

constexpr int NBUCKET = 16;  // chunk length classes of the tabulation (512 terms each)

// Lanes that tabulate chunk c, one per start value: 2^P (0 for a chunk of
// zeros, a chunk whose start base leaves the binade, or P > 6), and its
// length class.
__device__ __forceinline__ int chunk_lanes(const Peak* __restrict__ pkq, long nc, long npad,
                                           long c, int& bucket) {
    const Chunk k = chunk_of(pkq, nc, npad, c);
    const long len = k.last - k.first + 1;
    bucket = len >> 9 < NBUCKET - 1 ? (int)(len >> 9) : NBUCKET - 1;
    if (k.zero) return 0;
    const StitchInfo i = stitch_info(k);
    return i.ok ? 1 << i.P : 0;
}

// Per sequence (blockIdx.x): the list of tabulation lanes, one per start
// value base + r of each tabulated chunk, entry c * 64 + r, the chunks
// ordered from
// the longest length class to the shortest (so that a wavefront's lanes run
// about as long as each other, and the longest start first), at
// lanes[q * 64 nc ...]; their number at counts[q].
__global__ void __launch_bounds__(1024) pk_class_kernel(const Peak* __restrict__ pk, long npad,
                                                        int* __restrict__ lanes,
                                                        int* __restrict__ counts) {
    __shared__ int sc[NBUCKET][1024];
    __shared__ int base[NBUCKET + 1];
    const long nc = npad / SEQ_CBLK;
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    const Peak* pkq = pk + (long)q * nc;
    const long per = (nc + 1023) / 1024;
    const long b = (long)tid * per;
    const long e = b + per < nc ? b + per : nc;
    int cnt[NBUCKET];
#pragma unroll
    for (int k = 0; k < NBUCKET; ++k) cnt[k] = 0;
    for (long c = b; c < e; ++c) {
        int bk;
        const int L = chunk_lanes(pkq, nc, npad, c, bk);
#pragma unroll
        for (int k = 0; k < NBUCKET; ++k) cnt[k] += k == bk ? L : 0;
    }
#pragma unroll
    for (int k = 0; k < NBUCKET; ++k) sc[k][tid] = cnt[k];
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int add[NBUCKET];
#pragma unroll
        for (int k = 0; k < NBUCKET; ++k) add[k] = tid >= o ? sc[k][tid - o] : 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NBUCKET; ++k) sc[k][tid] += add[k];
        __syncthreads();
    }
    if (tid == 0) {  // the longest class first
        int o = 0;
        for (int k = NBUCKET - 1; k >= 0; --k) {
            base[k] = o;
            o += sc[k][1023];
        }
        base[NBUCKET] = o;
        counts[q] = o;
    }
    __syncthreads();
    int off[NBUCKET];
#pragma unroll
    for (int k = 0; k < NBUCKET; ++k) off[k] = base[k] + sc[k][tid] - cnt[k];
    int* lq = lanes + (long)q * TAB * nc;
    for (long c = b; c < e; ++c) {
        int bk;
        const int L = chunk_lanes(pkq, nc, npad, c, bk);
        int o = 0;
#pragma unroll
        for (int k = 0; k < NBUCKET; ++k) {
            o = k == bk ? off[k] : o;
            off[k] += k == bk ? L : 0;
        }
        for (int r = 0; r < L; ++r) lq[o + r] = (int)(c * TAB + r);
    }
}

// Adds x to each of the NCH sums s[r].  With CERT, also keeps for each the
// highest exponent field met (em, the field in bits 20..30) and the
// smallest distance mm of a sum to the edges of its binade [2^e, 2^(e+1))
// (field 0: [0, 2^-1022)); both differences are exact.
template <int NCH, bool CERT>
__device__ __forceinline__ void add_all(double x, double (&s)[NCH], double (&mm)[NCH],
                                        unsigned (&em)[NCH]) {
#pragma unroll
    for (int r = 0; r < NCH; ++r) {
        s[r] = s[r] + x;
        if (CERT) {
            const unsigned e = (unsigned)(ubits(s[r]) >> 32) & 0x7ff00000u;
            em[r] = e > em[r] ? e : em[r];
            const double lo = dfrom((unsigned long long)e << 32);
            const double up = dfrom((unsigned long long)(e + 0x00100000u) << 32);
            const double a = fabs(s[r]);
            // (a NaN here comes with em = 0x7ff, which voids the certificate)
            mm[r] = fmin(mm[r], fmin(a - lo, up - a));
        }
    }
}

// Each lane adds its own len terms at p to its NCH sums, in order (len = 0:
// nothing; a sum is never -0, so the +0 fed to a lane past its end leaves it
// as it is).  32 terms are in flight while the previous 32 are added.
// Reads reach p[max(len, 1) + 31]: the term arrays are followed by at least
// that much readable memory (the workspace).
template <int NCH, bool CERT>
__device__ __forceinline__ void run_terms(const double* __restrict__ p, long len,
                                          double (&s)[NCH], double (&mm)[NCH],
                                          unsigned (&em)[NCH]) {
    constexpr int BT = 32;
    const long maxlen = wave_max_long(len);
    const long minlen = wave_min_long(len > 0 ? len : maxlen);
    double cur[BT], nxt[BT];
#pragma unroll
    for (int i = 0; i < BT; ++i) {
        cur[i] = 0.0;
        nxt[i] = 0.0;
    }
    if (len > 0) {
#pragma unroll
        for (int i = 0; i < BT; ++i) cur[i] = p[i];
    }
    long base = 0;
    for (; base + BT <= minlen; base += BT) {  // every lane with terms adds all 16
        if (len > 0) {
#pragma unroll
            for (int i = 0; i < BT; ++i) nxt[i] = p[base + BT + i];
        }
#pragma unroll
        for (int i = 0; i < BT; ++i) add_all<NCH, CERT>(cur[i], s, mm, em);
#pragma unroll
        for (int i = 0; i < BT; ++i) cur[i] = nxt[i];
    }
    for (; base < maxlen; base += BT) {  // the tails: +0 past a lane's end
        if (base + BT < len) {
#pragma unroll
            for (int i = 0; i < BT; ++i) nxt[i] = p[base + BT + i];
        }
#pragma unroll
        for (int i = 0; i < BT; ++i) add_all<NCH, CERT>(base + i < len ? cur[i] : 0.0, s, mm, em);
#pragma unroll
        for (int i = 0; i < BT; ++i) cur[i] = nxt[i];
    }
}

// The exactness certificate of a tabulated run from the start v0 (exponent
// field of key, residue bits P) whose sums kept the distance mm from their
// binades' edges and reached exponent field em at most: the largest |delta|
// for which the run from v0 + delta ulp(v0) (delta a multiple of 2^P) ends
// exactly delta ulp(v0) from the tabulated end, or -1 (only delta = 0).
// Proof: let w be an exact sum of the run from v0 and v_i = RN(w), in the
// binade of field e with spacing g, at distance m_i >= g from its edges; w
// then lies in the same binade, at least m_i - g/2 inside.  If e <= e0 + P - 1,
// D = delta ulp(v0) is an even multiple of g, so when |D| <= m_i - g, w + D
// rounds to v_i + D (rounding within a binade commutes with shifts by even
// multiples of the spacing, ties to even included); by induction the shifted
// run stays D from the tabulated one.  With g <= 2^(P-1) ulp(v0) for every i,
// |delta| <= floor(mm / ulp(v0)) - 2^(P-1) suffices.
__device__ __forceinline__ double certificate(double mm, unsigned em, unsigned key, int P) {
    const int e0 = (int)(key & 0x7ff);
    if ((int)(em >> 20) > e0 + P - 1) return -1.0;
    const double t = ldexp(mm, 1075 - e0);  // mm / ulp(v0), exact unless below 1
    const double allow = (t >= 1.0 ? floor(t) : 0.0) - (double)(1 << (P - 1));
    if (allow < 0.0) return -1.0;
    return allow < ALLOW_MAX ? allow : ALLOW_MAX;
}

// Tabulation for launches with few sums: lane i of sequence blockIdx.y takes
// entry i of the lane list and runs its start value through the chunk's
// terms from global memory, 32 in flight per lane.  With a handful of sums
// there are only a few hundred wavefronts, too few for the staged kernel
// below, whose stages put each wavefront's loads, LDS round trip and
// additions one after another.
__global__ void __launch_bounds__(256) pk_tab_lanes_kernel(const double* __restrict__ terms,
                                                           long npad, const Peak* __restrict__ pk,
                                                           const int* __restrict__ lanes,
                                                           const int* __restrict__ counts,
                                                           double* __restrict__ tab,
                                                           double* __restrict__ cert) {
    const long nc = npad / SEQ_CBLK;
    const int q = blockIdx.y;
    const int n = counts[q];
    const int* lq = lanes + (long)q * TAB * nc;
    for (long i0 = (long)blockIdx.x * 256; i0 < n; i0 += (long)gridDim.x * 256) {
        const long i = i0 + threadIdx.x;
        const bool active = i < n;
        if (__ballot(active) == 0ull) continue;  // uniform across the wavefront
        const int ent = active ? lq[i] : 0;
        const long c = ent / TAB;
        const int r = ent % TAB;
        const Chunk k = chunk_of(pk + (long)q * nc, nc, npad, c);
        const StitchInfo info = stitch_info(k);
        const long long I = info.base + r;
        double s[1] = {dfrom(((unsigned long long)info.key << 52) |
                             ((unsigned long long)(I < 0 ? -I : I) & MANT))};
        double mm[1] = {__builtin_huge_val()};
        unsigned em[1] = {0u};
        const long len = active ? k.last - k.first + 1 : 0;
        run_terms<1, true>(terms + (long)q * npad + (active ? k.first : 0), len, s, mm, em);
        if (active) {
            tab[((long)q * nc + c) * TAB + r] = s[0];
            cert[((long)q * nc + c) * TAB + r] = certificate(mm[0], em[0], info.key, info.P);
        }
    }
}

// Tabulation stages: each chunk of a wavefront is read 32 terms at a time
// into a row of the wavefront's LDS block by coalesced loads (two rows per
// load instruction), and every lane adds its chunk's row in order.  A
// wavefront spans at most 33 chunks: pk_class_kernel gives each chunk 2^P >=
// 2 consecutive lanes, so only the first and the last can be cut.
constexpr int TAB_STAGE = 32;                   // terms per row and stage
constexpr int TAB_ROWS = 33;                    // chunks per wavefront
constexpr int TAB_PAIRS = (TAB_ROWS + 1) / 2;   // load instructions per stage
constexpr int TAB_STRIDE = TAB_STAGE + 1;       // row stride (doubles): rows on distinct banks
constexpr int TAB_FEW_SUMS = 8;  // up to this many sums per launch: pk_tab_lanes_kernel

// Tabulation: lane i of sequence blockIdx.x takes entry i of the lane list,
// start value base + r of chunk c, and runs it through the chunk's terms,
// keeping its certificate (one wavefront per workgroup, blockIdx.x the
// sequence; blockIdx.y strides over the list).  The additions are those of run_terms, in the same order:
// a lane past its chunk's end adds +0, which leaves its sum as it is.
__global__ void __launch_bounds__(64) pk_tab_kernel(const double* __restrict__ terms, long npad,
                                                    const Peak* __restrict__ pk,
                                                    const int* __restrict__ lanes,
                                                    const int* __restrict__ counts,
                                                    double* __restrict__ tab,
                                                    double* __restrict__ cert) {
    __shared__ double s_rows[TAB_ROWS * TAB_STRIDE];
    __shared__ long s_first[TAB_ROWS], s_len[TAB_ROWS];
    const long nc = npad / SEQ_CBLK;
    const int q = blockIdx.x;
    const int n = counts[q];
    const int* lq = lanes + (long)q * TAB * nc;
    const double* tq = terms + (long)q * npad;
    const int lane = threadIdx.x;
    const int half = lane >> 5, col = lane & 31;
    for (long i0 = (long)blockIdx.y * 64; i0 < n; i0 += (long)gridDim.y * 64) {
        const long i = i0 + lane;
        const bool active = i < n;  // lane 0 always is
        const int ent = active ? lq[i] : 0;
        const long c = active ? ent / TAB : -1;
        const int r = ent % TAB;
        const Chunk k = chunk_of(pk + (long)q * nc, nc, npad, active ? c : 0);
        const StitchInfo info = stitch_info(k);
        const long long I = info.base + r;
        double s[1] = {dfrom(((unsigned long long)info.key << 52) |
                             ((unsigned long long)(I < 0 ? -I : I) & MANT))};
        double mm[1] = {__builtin_huge_val()};
        unsigned em[1] = {0u};
        const long len = active ? k.last - k.first + 1 : 0;
        // the wavefront's chunks, in list order: row d is the d-th chunk
        const long cprev = __shfl_up(c, 1, 64);
        const bool leader = active && (lane == 0 || cprev != c);
        const unsigned long long lmask = __ballot(leader);
        const int D = __popcll(lmask);
        if (D > TAB_ROWS) {  // not produced by pk_class_kernel; kept correct, term by term
            const double* pc = tq + (active ? k.first : 0);
            for (long t = 0; t < len; ++t) add_all<1, true>(pc[t], s, mm, em);
        } else {
            const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
            const int row = __popcll(lmask & upto) - 1;  // inactive lanes: the last row
            if (leader) {
                s_first[row] = k.first;
                s_len[row] = len;
            }
            __syncthreads();
            const long maxlen = wave_max_long(len);
            double pre[TAB_PAIRS];  // the next stage's terms, two rows per load
            auto load = [&](long base) {
#pragma unroll
                for (int p = 0; p < TAB_PAIRS; ++p) {
                    const int d = 2 * p + half;
                    double v = 0.0;
                    if (d < D) {
                        const long idx = base + col;
                        if (idx < s_len[d]) v = tq[s_first[d] + idx];
                    }
                    pre[p] = v;
                }
            };
            load(0);
            for (long base = 0; base < maxlen; base += TAB_STAGE) {
#pragma unroll
                for (int p = 0; p < TAB_PAIRS; ++p) {
                    const int d = 2 * p + half;
                    if (d < D) s_rows[d * TAB_STRIDE + col] = pre[p];
                }
                __syncthreads();
                if (base + TAB_STAGE < maxlen) load(base + TAB_STAGE);
                const double* rowp = s_rows + row * TAB_STRIDE;
#pragma unroll
                for (int t = 0; t < TAB_STAGE; ++t) add_all<1, true>(rowp[t], s, mm, em);
                __syncthreads();  // the row block is written again next stage
            }
        }
        if (active) {
            tab[((long)q * nc + c) * TAB + r] = s[0];
            cert[((long)q * nc + c) * TAB + r] = certificate(mm[0], em[0], info.key, info.P);
        }
    }
}

// s + x[first] + ... + x[first + len - 1] in order, first and len uniform
// across the wavefront: every lane runs the same chain on terms read
// through the scalar unit, the next 16 in flight while the current 16 are
// added.  Returns the same value in every lane.
__device__ __forceinline__ double run_wave(const double* __restrict__ x, long first, long len,
                                           double s) {
    const double* p = x + first;
    long i = 0;
    if (len >= 16) {
        double t[16], tn[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) t[k] = p[k];
        for (; i + 32 <= len; i += 16) {
#pragma unroll
            for (int k = 0; k < 16; ++k) tn[k] = p[i + 16 + k];
#pragma unroll
            for (int k = 0; k < 16; ++k) s = s + t[k];
#pragma unroll
            for (int k = 0; k < 16; ++k) t[k] = tn[k];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) s = s + t[k];
        i += 16;
    }
    for (; i < len; ++i) s = s + p[i];
    return s;
}

__device__ __forceinline__ long long imant_u(unsigned long long u) {
    const long long m = (long long)((u & MANT) | (1ull << 52));
    return (u >> 63) ? -m : m;
}
__device__ __forceinline__ bool nan_u(unsigned long long u) {
    return (u & (0x7ffull << 52)) == (0x7ffull << 52) && (u & MANT) != 0;
}

// The end of a chunk from a start of integer I and exponent field es in the
// chunk's predicted binade, given e = the end tabulated from the start value
// base + r (r = I mod 2^P) and that run's certificate allow: the integer of e
// plus delta = I - base - r scaled to e's binade (k = es - e's exponent
// field), as the bits of a value in e's binade.  False when delta is neither
// 0 nor within the certificate, e is not a normal value, the scaling would
// not be exact or could overflow, or the end leaves e's binade.
__device__ __forceinline__ bool chunk_end(long long I, int es, int P, long long base,
                                          unsigned long long e, long long allow,
                                          unsigned long long& out, int& k) {
    const long long r = I & ((1ll << P) - 1);
    const long long delta = I - base - r;  // a multiple of 2^P
    if (delta != 0 && (delta > allow || -delta > allow)) return false;
    const int ee = (int)((e >> 52) & 0x7ff);
    if (ee == 0 || ee == 0x7ff) return false;
    const long long IE = imant_u(e);
    k = es - ee;
    long long D;
    if (k >= 0) {
        if (k > 10 || delta >= (1ll << 50) || delta <= -(1ll << 50)) return false;
        D = delta * (1ll << k);
    } else {
        if (-k >= P) return false;
        D = delta >> (-k);
    }
    const long long Ie = IE + D;
    if ((Ie < 0) != (IE < 0)) return false;
    const long long a = Ie < 0 ? -Ie : Ie;
    if (a < (1ll << 52) || a >= (1ll << 53)) return false;
    out = (e & (0xfffull << 52)) | ((unsigned long long)a & MANT);
    return true;
}

// One chunk of a chunk-by-chunk stitch on the exact start u (uniform across
// the wavefront): Eg and Cg hold the chunk's tabulated ends and
// certificates, those of residue r in lane r.
__device__ __forceinline__ bool stitch_step(unsigned long long& u, unsigned key, int P,
                                            long long base, double Eg, double Cg) {
    if ((unsigned)(u >> 52) != key) return false;
    const long long I = imant_u(u);
    const int r = (int)(I & ((1ll << P) - 1));
    unsigned long long o;
    int k;
    if (!chunk_end(I, (int)(key & 0x7ff), P, base, ubits(rl_d(Eg, r)), (long long)rl_d(Cg, r), o,
                   k))
        return false;
    u = rfl_u(o);
    return true;
}

struct StitchCount {
    long long small, big, fall, zero;
};

// Chunks [cb, ce) of a sequence from the exact start u, one after another
// (the whole wavefront, the sum uniform across it): tabulated where the
// checks pass, term by term otherwise, a chunk of zeros passed through.  A
// chunk of zeros leaves the sum as it is: s + 0 and s - 0 are s for every s
// but -0 and NaN, and the sum, which starts at +0, is never -0.  Writes each
// chunk's start to sq[c] when sq is given; returns the end.  Lane r holds
// row r of the tables of the STITCH_G chunks in hand, so E[r] is a read of
// lane r; the next round's rows load meanwhile.
__device__ unsigned long long stitch_range(const double* __restrict__ x,
                                           const Peak* __restrict__ pkq,
                                           const double* __restrict__ tq,
                                           const double* __restrict__ cq, long nc, long npad,
                                           long cb, long ce, unsigned long long u,
                                           double* __restrict__ sq, StitchCount& ct) {
    const int lane = threadIdx.x & 63;
    double E[STITCH_G], En[STITCH_G], C[STITCH_G], Cn[STITCH_G];
#pragma unroll
    for (int g = 0; g < STITCH_G; ++g) {
        E[g] = cb + g < ce ? tq[(cb + g) * TAB + lane] : 0.0;
        C[g] = cb + g < ce ? cq[(cb + g) * TAB + lane] : 0.0;
    }
    Chunk mine = chunk_of(pkq, nc, npad, cb + lane < ce ? cb + lane : cb);
    for (long c0 = cb; c0 < ce; c0 += STITCH_G) {
        const int g_n = ce - c0 < STITCH_G ? (int)(ce - c0) : STITCH_G;
        const long c1 = c0 + STITCH_G;
#pragma unroll
        for (int g = 0; g < STITCH_G; ++g) {
            En[g] = c1 + g < ce ? tq[(c1 + g) * TAB + lane] : 0.0;
            Cn[g] = c1 + g < ce ? cq[(c1 + g) * TAB + lane] : 0.0;
        }
        const Chunk next = chunk_of(pkq, nc, npad, c1 + lane < ce ? c1 + lane : cb);
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
                if (P <= 3) ++ct.small;
                else if (P <= SEQ_PMAX) ++ct.big;
                if (zero && !nan_u(u)) {
                    ++ct.zero;
                } else if (!(ok && stitch_step(u, key, P, base, E[g], C[g]))) {
                    const long first = rl_l(mine.first, g);
                    const long last = rl_l(mine.last, g);
                    u = rfl_u(ubits(run_wave(x, first, last - first + 1,
                                             __longlong_as_double((long long)u))));
                    ++ct.fall;
                }
            }
        }
        if (sq && lane < g_n) sq[c0 + lane] = mystart;
#pragma unroll
        for (int g = 0; g < STITCH_G; ++g) {
            E[g] = En[g];
            C[g] = Cn[g];
        }
        mine = next;
    }
    return u;
}

// The map of a group of GROUP chunks, for start residue r mod 2^PG: from
// the start gbase + r (gbase = the integer of the first chunk's predicted
// start rounded down to a multiple of 2^PG, in its binade) the group ends at
// end, and from gbase + r + delta (delta a multiple of 2^PG) at
// end + delta 2^K exactly when lo <= delta <= hi.  valid = 0: no map for
// this residue.
struct GroupMap {
    unsigned long long end;
    long long lo, hi;
    int K, valid;
};
// Per group: gbase and its binade's key (ok = 0: no maps), and whether all
// its chunks are zeros.
struct GroupHead {
    long long base;
    unsigned key;
    short ok, zero;
};

// Level 1 of the stitch, one wavefront per group: lane r carries the start
// gbase + r through the group's chunks with the chunk-by-chunk arithmetic.
// A start that differs by delta (a multiple of 2^PG) follows the same
// residues while PG + K >= P at every chunk (delta 2^K is then a multiple of
// 2^P; K falls by one for each binade the sum rises), and its offset from
// each chunk's tabulated run, the lane's offset d plus delta 2^K, must stay
// within that run's certificate A: (-A - d) / 2^K <= delta <= (A - d) / 2^K
// for every chunk.  It then ends delta 2^K from the lane's end.  Chunks of
// zeros pass the value through.
__global__ void __launch_bounds__(64) pk_group_kernel(const double* __restrict__ tab,
                                                      const double* __restrict__ cert, long npad,
                                                      const Peak* __restrict__ pk,
                                                      GroupMap* __restrict__ gm,
                                                      GroupHead* __restrict__ gh) {
    const long nc = npad / SEQ_CBLK;
    const long ng = (nc + GROUP - 1) / GROUP;
    const long g = blockIdx.x;
    const int q = blockIdx.y;
    const int lane = threadIdx.x;
    const long cb = g * GROUP;
    const int cn = (int)(nc - cb < GROUP ? nc - cb : GROUP);
    const Peak* pkq = pk + (long)q * nc;
    const double* tq = tab + (long)q * nc * TAB;
    const double* cq = cert + (long)q * nc * TAB;
    const Chunk mine = chunk_of(pkq, nc, npad, cb + (lane < cn ? lane : 0));
    const StitchInfo info = stitch_info(mine);
    const bool allzero = __all(lane >= cn || mine.zero) != 0;
    // the group's base: chunk cb's predicted start (lane 0's chunk)
    const double V0 = rfl_d(mine.V);
    GroupHead head;
    head.key = key_of(V0);
    head.zero = allzero ? 1 : 0;
    head.ok = 0;
    head.base = 0;
    if (normal_finite(V0)) {
        const long long iv = imant(V0);
        const long long b = iv - (iv & ((1ll << PG) - 1));
        const long long e = b + ((1ll << PG) - 1);
        const long long ab = b < 0 ? -b : b, ae = e < 0 ? -e : e;
        head.ok = ab >= (1ll << 52) && ab < (1ll << 53) && ae >= (1ll << 52) && ae < (1ll << 53);
        head.base = b;
    }
    GroupMap* out = gm + ((long)q * ng + g) * NRES;
    if (lane == 0) gh[(long)q * ng + g] = head;
    GroupMap res;
    res.end = 0;
    res.lo = 0;
    res.hi = -1;
    res.K = 0;
    res.valid = 0;
    if (allzero || !head.ok) {  // uniform across the wavefront
        out[lane] = res;
        return;
    }
    bool valid = true;
    unsigned long long u;
    {
        const long long I = head.base + lane;
        const unsigned long long a = (unsigned long long)(I < 0 ? -I : I);
        u = ((unsigned long long)head.key << 52) | (a & MANT);
    }
    long long lo = -(1ll << 62), hi = 1ll << 62;
    int K = 0;
    double E[STITCH_G], En[STITCH_G], C[STITCH_G], Cn[STITCH_G];
#pragma unroll
    for (int t = 0; t < STITCH_G; ++t) {
        E[t] = t < cn ? tq[(cb + t) * TAB + lane] : 0.0;
        C[t] = t < cn ? cq[(cb + t) * TAB + lane] : 0.0;
    }
    for (int t0 = 0; t0 < cn; t0 += STITCH_G) {
        const int t1 = t0 + STITCH_G;
#pragma unroll
        for (int t = 0; t < STITCH_G; ++t) {
            En[t] = t1 + t < cn ? tq[(cb + t1 + t) * TAB + lane] : 0.0;
            Cn[t] = t1 + t < cn ? cq[(cb + t1 + t) * TAB + lane] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < STITCH_G; ++t) {
            const int idx = t0 + t;
            if (idx < cn && !__builtin_amdgcn_readlane(info.zero, idx)) {
                const bool okc = __builtin_amdgcn_readlane(info.ok, idx) != 0;
                const int P = okc ? __builtin_amdgcn_readlane(info.P, idx) : 0;
                const unsigned key = (unsigned)__builtin_amdgcn_readlane((int)info.key, idx);
                const long long base = rl_l(info.base, idx);
                const long long I = imant_u(u);
                const int r = valid ? (int)(I & ((1ll << P) - 1)) : 0;
                const double Ev = __shfl(E[t], r, 64);
                const long long allow = (long long)__shfl(C[t], r, 64);
                if (valid) {
                    unsigned long long o = 0;
                    int k = 0;
                    valid = okc && (unsigned)(u >> 52) == key && PG + K >= P &&
                            chunk_end(I, (int)(key & 0x7ff), P, base, ubits(Ev), allow, o, k);
                    if (valid) {
                        const long long A = allow > 0 ? allow : 0;
                        const long long d = I - base - (I & ((1ll << P) - 1));
                        const long long a = -A - d, b = A - d;
                        long long l, h;
                        if (K >= 0) {
                            const int sh = K < 62 ? K : 62;
                            l = -((-a) >> sh);  // ceil(a / 2^K)
                            h = b >> sh;        // floor(b / 2^K)
                        } else {
                            l = a * (1ll << (-K));
                            h = b * (1ll << (-K));
                        }
                        lo = l > lo ? l : lo;
                        hi = h < hi ? h : hi;
                        u = o;
                        K += k;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < STITCH_G; ++t) {
            E[t] = En[t];
            C[t] = Cn[t];
        }
    }
    if (valid && lo <= hi) {
        res.end = u;
        res.lo = lo;
        res.hi = hi;
        res.K = K;
        res.valid = 1;
    }
    out[lane] = res;
}

// Level 2, one wavefront per sequence: the exact running sum from group to
// group, through the group's map where the start lies in its first chunk's
// binade and within the map's range, chunk by chunk otherwise.  Writes the
// group starts and the sum.  The maps pass through LDS in batches of
// MAP_BATCH groups (the next batch loads while one is walked); the launch
// asks for STITCH_LDS_BYTES, which also keeps every sequence on a CU of its
// own (serial chains of scalar work that would otherwise share a CU's
// scalar unit).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
pk_stitch_kernel(const double* __restrict__ terms, long npad,
                                                       const Peak* __restrict__ pk,
                                                       const double* __restrict__ tab,
                                                       const double* __restrict__ cert,
                                                       const GroupMap* __restrict__ gm,
                                                       const GroupHead* __restrict__ gh,
                                                       double* __restrict__ gstart,
                                                       double* __restrict__ starts,
                                                       unsigned long long* __restrict__ flags,
                                                       double* __restrict__ sums,
                                                       long long* __restrict__ stats) {
    extern __shared__ GroupMap ring[];  // [2][MAP_BATCH][NRES]
    const long nc = npad / SEQ_CBLK;
    const long ng = (nc + GROUP - 1) / GROUP;
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const double* x = terms + (long)q * npad;
    const Peak* pkq = pk + (long)q * nc;
    const double* tq = tab + (long)q * nc * TAB;
    const double* cq = cert + (long)q * nc * TAB;
    const GroupMap* gmq = gm + (long)q * ng * NRES;
    const GroupHead* ghq = gh + (long)q * ng;
    double* gs = gstart + (long)q * (ng + 1);
    if (lane == 0) flags[q] = 0ull;
    const bool want = stats != nullptr;
    const long long t0 = want ? (long long)clock64() : 0;
    StitchCount ct{0, 0, 0, 0};
    long long n_walk = 0;
    // the head: the terms up to the first block's peak, from 0 (the same
    // value in every lane; readfirstlane tells the compiler so, which keeps
    // the walk's integer work on the scalar unit)
    unsigned long long u = rfl_u(ubits(run_wave(x, 0, pkq[0].p + 1, 0.0)));
    // the next batch's maps, staged in registers (lane r: residue r of each
    // group; eight named registers, not an array, so that none lands in
    // scratch memory)
    static_assert(MAP_BATCH == 8, "the staging below names eight groups");
    const double4* gm4 = reinterpret_cast<const double4*>(gmq);
    double4 v0, v1, v2, v3, v4, v5, v6, v7;
#define HYD_MAPS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
#define HYD_LOAD(b) v##b = gm4[(G1 + b < ng ? G1 + b : ng - 1) * NRES + lane];
#define HYD_STORE(b) buf4[b * NRES + lane] = v##b;
    GroupHead hv;
    {
        const long G1 = 0;
        HYD_MAPS(HYD_LOAD)
    }
    hv = ghq[lane < MAP_BATCH && lane < ng ? lane : 0];
    for (long G0 = 0; G0 < ng; G0 += MAP_BATCH) {
        GroupMap* buf = ring + ((G0 / MAP_BATCH) & 1) * MAP_BATCH * NRES;
        double4* buf4 = reinterpret_cast<double4*>(buf);
        HYD_MAPS(HYD_STORE)
        const GroupHead hd = hv;  // lane b: the head of group G0 + b
        __syncthreads();
        const long G1 = G0 + MAP_BATCH;
        HYD_MAPS(HYD_LOAD)
        hv = ghq[lane < MAP_BATCH && G1 + lane < ng ? G1 + lane : 0];
        const int bn = ng - G0 < MAP_BATCH ? (int)(ng - G0) : MAP_BATCH;
        for (int b = 0; b < bn; ++b) {
            const long g = G0 + b;
            if (lane == 0) gs[g] = __longlong_as_double((long long)u);
            if (__builtin_amdgcn_readlane((int)hd.zero, b) && !nan_u(u)) continue;
            bool done = false;
            const bool ok0 = __builtin_amdgcn_readlane((int)hd.ok, b) != 0;
            const unsigned key0 = (unsigned)__builtin_amdgcn_readlane((int)hd.key, b);
            if (ok0 && (unsigned)(u >> 52) == key0) {
                const long long I = imant_u(u);
                const int r = (int)(I & ((1ll << PG) - 1));
                const long long delta = I - rl_l(hd.base, b) - r;
                const GroupMap mp = buf[b * NRES + r];  // the same address in every lane
                if (__builtin_amdgcn_readfirstlane(mp.valid) &&
                    delta >= (long long)rfl_u((unsigned long long)mp.lo) &&
                    delta <= (long long)rfl_u((unsigned long long)mp.hi)) {
                    const unsigned long long A = rfl_u(mp.end);
                    const int K = __builtin_amdgcn_readfirstlane(mp.K);
                    const long long Ie =
                        imant_u(A) + (K >= 0 ? delta * (1ll << K) : delta >> (-K));
                    const long long aa = Ie < 0 ? -Ie : Ie;
                    u = rfl_u((A & (0xfffull << 52)) | ((unsigned long long)aa & MANT));
                    done = true;
                }
            }
            if (!done) {
                ++n_walk;
                const long c0 = g * GROUP;
                u = stitch_range(x, pkq, tq, cq, nc, npad, c0, c0 + GROUP < nc ? c0 + GROUP : nc, u,
                                 nullptr, ct);
            }
        }
        __syncthreads();
    }
#undef HYD_STORE
#undef HYD_LOAD
#undef HYD_MAPS
    const double s = __longlong_as_double((long long)u);
    if (lane == 0) {
        gs[ng] = s;
        starts[(long)q * (nc + 1) + nc] = s;
        sums[2 * q] = s;
        sums[2 * q + 1] = 0.0;
        if (want) {
            long long* o = stats + (long)SEQ_STATS * q;
            o[0] = nc;
            o[1] = 0;
            o[2] = 0;
            o[3] = 0;
            o[4] = 0;
            o[5] = (long long)clock64() - t0;
            o[6] = n_walk;
            o[7] = 0;
            o[8] = 0;
        }
    }
}

// Level 3, one wavefront per group: the exact start of every chunk, chunk by
// chunk from the group's exact start, and the chunk counters.
__global__ void __launch_bounds__(64) pk_fill_kernel(const double* __restrict__ terms, long npad,
                                                     const Peak* __restrict__ pk,
                                                     const double* __restrict__ tab,
                                                     const double* __restrict__ cert,
                                                     const double* __restrict__ gstart,
                                                     double* __restrict__ starts,
                                                     long long* __restrict__ stats) {
    const long nc = npad / SEQ_CBLK;
    const long ng = (nc + GROUP - 1) / GROUP;
    const long g = blockIdx.x;
    const int q = blockIdx.y;
    const long cb = g * GROUP;
    const long ce = cb + GROUP < nc ? cb + GROUP : nc;
    StitchCount ct{0, 0, 0, 0};
    const unsigned long long u = rfl_u(ubits(gstart[(long)q * (ng + 1) + g]));
    (void)stitch_range(terms + (long)q * npad, pk + (long)q * nc, tab + (long)q * nc * TAB,
                       cert + (long)q * nc * TAB, nc, npad, cb, ce, u,
                       starts + (long)q * (nc + 1), ct);
    if (stats && threadIdx.x == 0) {
        unsigned long long* o = (unsigned long long*)(stats + (long)SEQ_STATS * q);
        atomicAdd(o + 1, (unsigned long long)ct.small);
        atomicAdd(o + 2, (unsigned long long)ct.big);
        atomicAdd(o + 3, (unsigned long long)ct.fall);
        atomicAdd(o + 8, (unsigned long long)ct.zero);
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
    if (wave_max_long(len) == 0) return;  // uniform across the wavefront
    double s[1] = {c < nc ? sq[c] : 0.0}, mm[1] = {0.0};
    unsigned em[1] = {0u};
    run_terms<1, false>(terms + (long)q * npad + (c < nc ? k.first : 0), len, s, mm, em);
    if (c < nc && ubits(s[0]) != ubits(sq[c + 1])) {
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

// Workspace of the peak-chunk methods, offsets in doubles.
struct PeaksLayout {
    long bs, pk, lanes, counts, tab, cert, starts, gstart, gm, gh, flags, total;
};
PeaksLayout peaks_layout(long npad, int nq) {
    const long nb = npad / SEQ_BLOCK1, nc = npad / SEQ_CBLK, ng = (nc + GROUP - 1) / GROUP;
    static_assert(sizeof(BlockStat) == 6 * sizeof(double), "BlockStat layout");
    static_assert(sizeof(Peak) == 2 * sizeof(double), "Peak layout");
    static_assert(sizeof(GroupMap) == 4 * sizeof(double), "GroupMap layout");
    static_assert(2 * MAP_BATCH * NRES * sizeof(GroupMap) <= STITCH_LDS_BYTES, "map ring");
    static_assert(sizeof(GroupHead) == 2 * sizeof(double), "GroupHead layout");
    PeaksLayout L;
    long o = 0;
    L.bs = o;
    o += 6 * nb * nq;
    L.pk = o;
    o += 2 * nc * nq;
    L.lanes = o;
    o += (TAB * nc * nq + 1) / 2;
    L.counts = o;
    o += (nq + 1) / 2;
    L.tab = o;
    o += (long)TAB * nc * nq;
    L.cert = o;
    o += (long)TAB * nc * nq;
    L.starts = o;
    o += (nc + 1) * nq;
    L.gstart = o;
    o += (ng + 1) * nq;
    L.gm = o;
    o += 4L * NRES * ng * nq;
    L.gh = o;
    o += 2 * ng * nq;
    L.flags = o;
    o += nq;
    L.total = o;
    return L;
}

}  // namespace

long seqsum_npad(long n) { return (n + SEQ_BLOCK2 - 1) / SEQ_BLOCK2 * SEQ_BLOCK2; }

long seqsum_workspace_len(long npad, int nq) { return peaks_layout(npad, nq).total; }

hipError_t launch_seqsum(int method, const double* terms, long npad, int nq, double* ws,
                         double* sums, hipStream_t st, long long* stats) {
    if (npad <= 0 || npad % SEQ_BLOCK2 != 0 || nq <= 0) return hipErrorInvalidValue;
    if (method == SEQSUM_CHAIN) {
        hipLaunchKernelGGL(seq_chain_kernel, dim3((unsigned)nq), dim3(256), 0, st, terms, npad,
                           sums);
        return hipGetLastError();
    }
    if (method != SEQSUM_PEAKS && method != SEQSUM_PEAKS_CHECKED && method != SEQSUM_PEAKS_STAGED)
        return hipErrorInvalidValue;
    // the chunk runs read up to 32 doubles past a sequence: the workspace
    // must follow the terms
    if (ws < terms + (long)nq * npad) return hipErrorInvalidValue;
    const long nb = npad / SEQ_BLOCK1, nc = npad / SEQ_CBLK, ng = (nc + GROUP - 1) / GROUP;
    const PeaksLayout L = peaks_layout(npad, nq);
    BlockStat* bs = reinterpret_cast<BlockStat*>(ws + L.bs);
    Peak* pk = reinterpret_cast<Peak*>(ws + L.pk);
    int* lanes = reinterpret_cast<int*>(ws + L.lanes);
    int* counts = reinterpret_cast<int*>(ws + L.counts);
    double* tab = ws + L.tab;
    double* cert = ws + L.cert;
    double* starts = ws + L.starts;
    double* gstart = ws + L.gstart;
    GroupMap* gm = reinterpret_cast<GroupMap*>(ws + L.gm);
    GroupHead* gh = reinterpret_cast<GroupHead*>(ws + L.gh);
    unsigned long long* flags = reinterpret_cast<unsigned long long*>(ws + L.flags);
    const unsigned uq = (unsigned)nq;
    hipLaunchKernelGGL(pk_block_kernel, dim3((unsigned)(nb / 4), uq), dim3(256), 0, st, terms,
                       npad, bs);
    hipLaunchKernelGGL(pk_scan_kernel, dim3(uq), dim3(1024), 0, st, bs, nb, pk);
    hipLaunchKernelGGL(pk_class_kernel, dim3(uq), dim3(1024), 0, st, pk, npad, lanes, counts);
    // about 2 lanes per chunk (most have P = 1); the grid strides over more
    if (nq <= TAB_FEW_SUMS && method != SEQSUM_PEAKS_STAGED) {
        hipLaunchKernelGGL(pk_tab_lanes_kernel, dim3((unsigned)((4 * nc + 255) / 256), uq),
                           dim3(256), 0, st, terms, npad, pk, lanes, counts, tab, cert);
    } else {
        // the sequence is the fastest launch index, so the wavefronts of every
        // sequence's longest chunks start first and the short ones fill the tail
        hipLaunchKernelGGL(pk_tab_kernel, dim3(uq, (unsigned)((4 * nc + 63) / 64)), dim3(64), 0,
                           st, terms, npad, pk, lanes, counts, tab, cert);
    }
    hipLaunchKernelGGL(pk_group_kernel, dim3((unsigned)ng, uq), dim3(64), 0, st, tab, cert, npad,
                       pk, gm, gh);
    hipLaunchKernelGGL(pk_stitch_kernel, dim3(uq), dim3(64), STITCH_LDS_BYTES, st, terms, npad,
                       pk, tab, cert, gm, gh, gstart, starts, flags, sums, stats);
    if (method == SEQSUM_PEAKS_CHECKED || stats) {
        hipLaunchKernelGGL(pk_fill_kernel, dim3((unsigned)ng, uq), dim3(64), 0, st, terms, npad,
                           pk, tab, cert, gstart, starts, stats);
    }
    if (method == SEQSUM_PEAKS_CHECKED) {
        hipLaunchKernelGGL(pk_verify_kernel, dim3((unsigned)((nc + 255) / 256), uq), dim3(256), 0,
                           st, terms, npad, pk, starts, flags, stats);
        hipLaunchKernelGGL(pk_redo_kernel, dim3(uq), dim3(256), 0, st, terms, npad, flags, sums,
                           stats);
    }
    return hipGetLastError();
}

}  // namespace hyd
