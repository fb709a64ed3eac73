// seqsum.hpp -- the sum s = ((0 + x[0]) + x[1]) + ... of n doubles, rounded
// after every addition exactly as a loop adding them one after another, for
// nq sequences at once.  The reference adds its multipole moment terms that
// way (gravity_solver_3d.c:182-198, 226-258).
//
// Three methods with the same result bits:
//   SEQSUM_CHAIN  one wavefront per sequence adds the terms in order.
//   SEQSUM_PEAKS  chunked with tabulated increments.  Write a double s of
//                 binade e (sign and exponent field) as s = I u, I a signed
//                 integer with 2^52 <= |I| < 2^53, u = ulp(s).  An addition
//                 rounds s + x to the grid 2^m u of the result's binade;
//                 while the result stays in that binade, adding an even
//                 multiple of 2^m u to s shifts the result by exactly as
//                 much (ties to even included), so RN(s + x) - s depends on s
//                 only through I mod 2^(m+1).  Over additions that follow one
//                 path of binades the increment depends only on I mod 2^P,
//                 P = 1 + how far the path rises above the start binade.
//                 Chunks therefore start at the running sum's predicted peak
//                 in each 4096-term block (from double-double block sums)
//                 and run to the next block's peak; each chunk is run from
//                 2^P start values near its prediction, one per residue
//                 (P <= 6), keeping how close each partial sum came to the
//                 edges of its binade.  That distance certifies the run: any
//                 start of the same residue within it follows the same
//                 binades and ends shifted by exactly its offset.  The exact
//                 sum is carried across a chunk in integers,
//                 end = E[I mod 2^P] + (start - variant start), where the
//                 offset is within the certificate (a chunk where it is not
//                 is added term by term; a chunk of zeros leaves the sum as
//                 it is).  The carrying is done 16 chunks at a time: each
//                 group of chunks is carried from 64 start values (residues
//                 mod 2^6) in parallel, which maps any start of the same
//                 residue within the certificates to its end, and one
//                 wavefront per sequence walks from group to group.
//   SEQSUM_PEAKS_CHECKED  SEQSUM_PEAKS, then every chunk is run again from
//                 its exact start, all in parallel, and must land on the next
//                 chunk's start; if any does not, the sequence is redone by
//                 the chain (a check of the certificates; tests use it).
//   SEQSUM_PEAKS_STAGED  SEQSUM_PEAKS with the LDS-staged tabulation whatever
//                 the number of sums (SEQSUM_PEAKS takes it above 8 sums per
//                 launch), so that tests reach it with single sums.
#pragma once

#include <hip/hip_runtime.h>

namespace hyd {

enum SeqSumMethod {
    SEQSUM_CHAIN = 0,
    SEQSUM_PEAKS = 1,
    SEQSUM_PEAKS_CHECKED = 2,
    SEQSUM_PEAKS_STAGED = 3
};

constexpr int SEQ_CHUNK = 16;                    // terms per lane in a block pass
constexpr int SEQ_BLOCK1 = 64 * SEQ_CHUNK;        // 1024 terms (one wavefront)
constexpr int SEQ_CBLK = 4 * SEQ_BLOCK1;          // 4096 terms: one chunk starts in each
constexpr int SEQ_BLOCK2 = 16 * SEQ_BLOCK1;       // term arrays are padded to this
constexpr int SEQ_PMAX = 6;                      // residue bits tabulated per chunk
constexpr int SEQ_STATS = 9;                     // counters per sequence

// Length of the term arrays: n rounded up to a multiple of SEQ_BLOCK2, the
// tail padded with zeros by the caller.
long seqsum_npad(long n);
// Workspace (doubles) of launch_seqsum for nq sequences of npad terms.
long seqsum_workspace_len(long npad, int nq);
// sums[2q] = the sequential sum of terms[q * npad .. + npad), sums[2q+1] = 0.
// stats (device, SEQ_STATS per sequence, or nullptr) for the peak-chunk
// methods: chunks; chunks with P = 1, 2 or 3; with P = 4..6; chunks added
// term by term; chunks whose check failed (SEQSUM_PEAKS_CHECKED); clock
// cycles of the walk from group to group; groups it walked chunk by chunk;
// 1 when the chain redid the sum; chunks of zeros passed through.
hipError_t launch_seqsum(int method, const double* terms, long npad, int nq, double* ws,
                         double* sums, hipStream_t st, long long* stats = nullptr);

}  // namespace hyd
