// nn_kernels.hpp -- launch interface of the learned initial-guess network.
//
// Activations are channels-last BF16: element (voxel v, channel c) at
// v * cp + c, with the channel count padded to cp = round_up(c, 4) so a
// 4-channel group is one 8-byte load.  Convolutions are 3x3x3, stride 1,
// zero padded, computed as implicit GEMMs on v_mfma_f32_16x16x16bf16_1k:
// M = voxels, N = output channels, K = 27 taps x cp input channels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hyd {
namespace nn {

inline int pad4(int c) { return (c + 3) & ~3; }
inline int ksteps_for(int cinp) { return (27 * cinp + 15) / 16; }
inline int ntiles_for(int cout) { return (cout + 15) / 16; }
// packed B fragments of one layer: [ntile][kstep][64 lanes][4] bf16
inline long packed_len(int cinp, int cout) {
    return (long)ntiles_for(cout) * ksteps_for(cinp) * 64 * 4;
}

struct ConvArgs {
    const uint16_t* x;     // input [nin^3][cinp]
    const uint16_t* wp;    // packed weights
    const float* bias;     // [cout] or nullptr
    const uint16_t* skip;  // [n^3][coutp] added after the ReLU, or nullptr
    uint16_t* yr;          // [n^3][coutp] relu(conv + bias) (or conv + bias), or nullptr
    uint16_t* yd;          // [n^3][coutp] yr + skip (with skip), or nullptr
    float* yf;             // [n^3] FP32 output of channel 0 (final layer), or nullptr
    const float* add;      // [n^3] added to yf (the monopole term), or nullptr
    int n;                 // output grid edge
    int up;                // 1: input lives on the n/2 grid (nearest upsampling fused)
    int cinp, cout, coutp;
    int relu;
    // [n^3][coutp] forward activation: yr is zeroed where it is <= 0 (the
    // ReLU derivative applied to an input gradient), or nullptr
    const uint16_t* mask = nullptr;
};

struct WgradArgs {
    const uint16_t* x;     // layer input [nin^3][cinp]
    const uint16_t* g;     // gradient w.r.t. the pre-activation [n^3][coutp]
    float* dw;             // [27][cinp][cout] FP32, accumulated
    float* db;             // [cout] FP32, accumulated
    float* part;           // scratch: wgrad_scratch_len() floats of block partials
    int n, up, cinp, cout, coutp;
};

// Blocks of the weight-gradient kernel and the scratch it needs.
constexpr int WGRAD_MAX_BLOCKS = 512;
inline long wgrad_part_len(int cinp, int cout) { return 27L * cinp * cout + cout; }
inline long wgrad_scratch_len() { return WGRAD_MAX_BLOCKS * wgrad_part_len(24, 24); }

hipError_t conv_forward(const ConvArgs& a, hipStream_t st);
hipError_t conv_wgrad(const WgradArgs& a, hipStream_t st);
// pack FP32 weights [27][cin][cout] into B fragments; transposed = 1 packs the
// adjoint (flipped taps, in/out swapped) used for the input gradient
hipError_t pack_weights(const float* w, int cin, int cout, int transposed, uint16_t* wp,
                        hipStream_t st);

// element-wise pieces (counts are voxels; cp = padded channels)
hipError_t avgpool2(const uint16_t* x, int n, int cp, uint16_t* y, hipStream_t st);
hipError_t sumpool2_add(const uint16_t* g, int n, int cp, uint16_t* out, int accumulate,
                        hipStream_t st);  // out[n/2] (+)= sum of 8 children
hipError_t unpool_add(const uint16_t* g, int n, int cp, uint16_t* out,
                      hipStream_t st);    // out[n] += g[n/2] / 8
hipError_t relu_mask(const uint16_t* g, const uint16_t* r, long count, uint16_t* out,
                     hipStream_t st);     // out = g * (r > 0)
hipError_t add_inplace(uint16_t* a, const uint16_t* b, long count, hipStream_t st);
hipError_t copy_bf16(const uint16_t* a, long count, uint16_t* out, hipStream_t st);
// sums of log10(max(rho, 1e-12)) and of its square into acc[0], acc[1]
hipError_t log_stats(const float* rho, long nvox, float* acc, hipStream_t st);
hipError_t log_stats_f64(const double* rho, long nvox, float* acc, hipStream_t st);
// Input features of an n^3 grid on the box [x0, x0 + L]^3 (cp = 4 layout):
// channel 0 = (log10(max(rho, 1e-12)) - mu) / sigma with mu, sigma from the
// sums in acc; channel 1 = 0.1 * mono, where mono = -1 / sqrt(r^2 + eps^2)
// is the softened monopole potential in units of G M / L about the centre
// of mass mom[1..3] (r in units of L, eps = 1.5 / n), also written to mono.
hipError_t make_features(const float* rho, int n, const float* acc, const double* mom,
                         double x0, double y0, double z0, double L, uint16_t* x, float* mono,
                         hipStream_t st);
hipError_t make_features_f64(const double* rho, int n, const float* acc, const double* mom,
                             double x0, double y0, double z0, double L, uint16_t* x, float* mono,
                             hipStream_t st);
// loss: mean squared error over all voxels of channel 0; writes gradient
// (bf16, cp = 4 layout) and accumulates the loss
hipError_t mse_loss(const float* pred, const float* target, long nvox, float scale, uint16_t* g,
                    float* loss, hipStream_t st);
hipError_t adam_step(float* w, const float* g, float* m, float* v, long count, float lr,
                     float b1, float b2, float eps, float c1, float c2, hipStream_t st);

// 2x coarsening of a density (mean of 8 cells, mass preserving)
hipError_t restrict2_f64(const double* fine, int n, double* coarse, hipStream_t st);
// Trilinear interpolation of a cell-centred nc^3 FP32 field onto the n^3
// grid (n a multiple of nc, edges clamped), times G M / L with M = mom[0]
// read on the device: the network output in physical units.
hipError_t prolong_to_f64(const float* coarse, int nc, int n, const double* mom, double G,
                          double L, double* fine, hipStream_t st);

// training densities on the device (see blob_field_kernel)
hipError_t blob_field(int n, const double* blob, int nblob, const double* mode, int nmode,
                      double floor_rho, double sigma, double* rho64, float* rho32,
                      hipStream_t st);
// out = phi * L / (G * M) with M = mom[0] read on the device
hipError_t scale_potential(const double* phi, long count, const double* mom, double G, double L,
                           float* out, hipStream_t st);

// Correction of a warm start (the correction network): the split FP32
// residual of an n^3 grid averaged over f^3 blocks into nc^3 = (n/f)^3;
// the features (r / s, -r / s) with s the root mean square (sumsq[0] = sum
// of r^2 in a fixed order through part, 256 doubles; mono zeroed); and the
// correction scale * s * P(e) on the interior, 0 on the boundary (FP32,
// split layout, into er / eb), P the separable Catmull-Rom interpolation
// from nc^3 to n^3 with odd reflection at the edges, through t1
// (nc * nc * n floats) and t2 (nc * n * n floats).
hipError_t corr_restrict(const float* rr, const float* rb, int n, int f, float* out,
                         hipStream_t st);
hipError_t corr_features(const float* r, int nc, double* sumsq, double* part, uint16_t* x,
                         float* mono, hipStream_t st);
hipError_t corr_apply(const float* e, int nc, int n, float* t1, float* t2, const double* sumsq,
                      double scale, float* er, float* eb, hipStream_t st);

}  // namespace nn
}  // namespace hyd
