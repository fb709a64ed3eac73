// seqsum.hpp -- the sum s = ((0 + x[0]) + x[1]) + ... of n doubles, rounded
// after every addition exactly as a loop adding them one after another, for
// nq sequences at once.  The reference adds its multipole moment terms that
// way (gravity_solver_3d.c:182-198, 226-258).
//
// Two methods with the same result bits:
//   SEQSUM_CHAIN      one wavefront per sequence adds the terms in order.
//   SEQSUM_COMPOSITE  while a running sum stays inside one binade
//                     [2^k, 2^(k+1)) (one sign) every addition rounds to the
//                     fixed grid u = 2^(k-52), so s + x rounds to
//                     s + u * rint(x / u) whenever the exact s + x stays in
//                     the binade, independently of s (rint: nearest integer;
//                     exact halves are left to the chain).  A parallel pass
//                     predicts each 16-term chunk's binade from approximate
//                     prefix sums and stores, for runs of chunks of one
//                     binade inside a 1024-term block (segments) and for
//                     blocks of 1024 and 16384 terms of one binade, the
//                     integer increment D and the range [lo, hi] of
//                     S_j + floor(x_j/u) relative to the entry S = s/u.  A
//                     walk then moves s block by block: a block applies when
//                     s lies in its binade and S + lo >= 2^52, S + hi < 2^53
//                     (every exact partial sum stayed in the binade);
//                     otherwise it opens the block, down to segments, and
//                     adds the terms of a segment that does not apply one by
//                     one.  A sequence with at least half of its 1024-blocks
//                     not of one binade (a sum that keeps crossing binades)
//                     is added by the chain instead.
#pragma once

#include <hip/hip_runtime.h>

namespace hyd {

enum SeqSumMethod { SEQSUM_CHAIN = 0, SEQSUM_COMPOSITE = 1 };

constexpr int SEQ_CHUNK = 16;                    // terms per chunk (one lane)
constexpr int SEQ_BLOCK1 = 64 * SEQ_CHUNK;        // 1024 terms (one wavefront)
constexpr int SEQ_BLOCK2 = 16 * SEQ_BLOCK1;       // 16384 terms

struct SeqDesc {
    long long D;   // increment of S over the block, in units of 2^(k-52)
    long long lo;  // min over terms of (increment so far + floor(x/u))
    long long hi;  // max of the same
    int k;         // binade of the running sum
    int flags;     // SEQ_VALID, SEQ_NEG (running sum negative), SEQ_ZERO
};
constexpr int SEQ_VALID = 1;
constexpr int SEQ_NEG = 2;
constexpr int SEQ_ZERO = 4;  // every term is +-0: the block leaves any s != -0 as it is
constexpr int SEQ_START = 8;  // a segment starts at this chunk (segment entries only)

// Length of the term arrays: n rounded up to a multiple of SEQ_BLOCK2, the
// tail padded with zeros by the caller.
long seqsum_npad(long n);
// Workspace (doubles) of launch_seqsum for nq sequences of npad terms.
long seqsum_workspace_len(long npad, int nq);
// sums[2q] = the sequential sum of terms[q * npad .. + npad), sums[2q+1] = 0.
// stats (device, 9 per sequence, or nullptr): what the composite walk did --
// blocks of 16384 and 1024, segments and chunks applied, chunks added term by
// term, 1024-blocks opened, clock cycles in all, in term-by-term adds and in
// opened 1024-blocks.
hipError_t launch_seqsum(int method, const double* terms, long npad, int nq, double* ws,
                         double* sums, hipStream_t st, long long* stats = nullptr);

}  // namespace hyd
