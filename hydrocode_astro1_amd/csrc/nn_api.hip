// nn_api.hip -- the learned initial guess for the Poisson solve: a 4-level
// 3D U-Net (3x3x3 convolutions, 65k parameters, BF16 weights and
// activations on MFMA, FP32 accumulation, FP32 master weights with Adam),
// trained on the device on density / potential pairs that this library's
// own solver produces.  No reference counterpart: the reference starts SOR
// from the previous potential or from zero (gravity_solver_3d.c:87, :566).
//
// Network (n = grid edge, channels in brackets):
//   x[2] -c1a-> e1a[8] -c1b-> e1[8] -pool-> p1 -c2a-> e2a[12] -c2b-> e2[12]
//   -pool-> p2 -c3a-> e3a[16] -c3b-> e3[16] -pool-> p3 -c4a-> e4a[24]
//   -c4b-> e4[24] -up,c5-> r5[16] (+e3 = d3) -up,c6-> r6[12] (+e2 = d2)
//   -up,c7-> r7[8] (+e1 = d1) -c8-> out[1] + mono
// ReLU after every convolution but the last; upsampling is nearest and
// fused into the next convolution's gather; skips add after the ReLU.
// Inputs: normalised log10 density and the softened monopole potential
// about the centre of mass; output: potential in units of G M / L, as the
// monopole plus a learned correction.
//
// The same network has a second kind, the start correction: its input is
// the residual r = f - L phi of a warm start averaged onto a grid of at most
// 64^3 (features r / s and -r / s, s the root mean square of r), and its
// output the correction e of the start (phi + e solves the problem) on that
// grid in units of C L^2 s, C a constant stored with the weights and L the
// box edge; e is interpolated to the solver's grid by cubic convolution
// (Catmull-Rom), odd about the boundary.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <random>
#include <vector>

#include "../../include/hydro_mi355x.h"
#include "host_common.hpp"
#include "nn_kernels.hpp"

hipStream_t hyd_stream();

namespace nnh {
using namespace hyd::nn;

struct LayerDef {
    int cin, cout, up, relu;
};
const LayerDef kLayers[] = {
    {2, 8, 0, 1},   {8, 8, 0, 1},   {8, 12, 0, 1},  {12, 12, 0, 1},
    {12, 16, 0, 1}, {16, 16, 0, 1}, {16, 24, 0, 1}, {24, 24, 0, 1},
    {24, 16, 1, 1}, {16, 12, 1, 1}, {12, 8, 1, 1},  {8, 1, 0, 0},
};
constexpr int kNL = 12;

long vox(int n) { return (long)n * n * n; }

// Activations (BF16, channels-last, padded to 4) and gradients of one grid.
struct Acts {
    int n = 0;
    uint16_t *x0 = nullptr, *e1a = nullptr, *e1 = nullptr, *p1 = nullptr, *e2a = nullptr,
             *e2 = nullptr, *p2 = nullptr, *e3a = nullptr, *e3 = nullptr, *p3 = nullptr,
             *e4a = nullptr, *e4 = nullptr, *r5 = nullptr, *d3 = nullptr, *r6 = nullptr,
             *d2 = nullptr, *r7 = nullptr, *d1 = nullptr;
    float *out = nullptr, *mono = nullptr;
    uint16_t *gout = nullptr, *gd1 = nullptr, *gd2 = nullptr, *gd3 = nullptr, *ge1 = nullptr,
             *ge2 = nullptr, *ge3 = nullptr, *ge4 = nullptr, *ge1a = nullptr, *ge2a = nullptr,
             *ge3a = nullptr, *ge4a = nullptr, *gp1 = nullptr, *gp2 = nullptr, *gp3 = nullptr,
             *s0 = nullptr, *s1 = nullptr;
    float* stats = nullptr;  // [2] sums of log10 rho and of its square
    float* loss = nullptr;   // [1]
    std::vector<void*> all;
};

struct Model {
    long woff[kNL], boff[kNL];  // offsets into params: weights [27][cinp][cout], bias [cout]
    long nparam = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    float* wg_part = nullptr;        // weight-gradient block partials
    uint16_t* wp[kNL] = {nullptr};   // packed B fragments, forward
    uint16_t* wpt[kNL] = {nullptr};  // packed B fragments, adjoint
    long step = 0;
    Acts acts;
    std::vector<double*> coarse;  // restricted densities of the guess
    std::vector<int> coarse_n;
    double* mom_tmp = nullptr;    // [4] moments given by the host (test entry)
    hipStream_t st = nullptr;
    int kind = 0;                 // 0 potential guess, 1 start correction
    double cscale = 1.0;          // C of the start correction
    double* sumsq = nullptr;      // [1] sum of squares of the correction input
    double* sumsq_part = nullptr; // [256] its partial sums
    float* rcoarse = nullptr;     // averaged residual of the start correction
    int rcoarse_n = 0;
    float *interp1 = nullptr, *interp2 = nullptr;  // interpolation stages of the correction
    int interp_n = 0;
};

void free_acts(Acts& a) {
    for (void* p : a.all) (void)hipFree(p);
    a.all.clear();
    a.n = 0;
}

int alloc_acts(Acts& a, int n) {
    if (a.n == n) return 0;
    free_acts(a);
    a.n = n;
    bool ok = true;
    auto al = [&](uint16_t** p, int level, int c) {
        const long cnt = vox(n >> level) * pad4(c);
        if (!ok || hipMalloc((void**)p, sizeof(uint16_t) * cnt) != hipSuccess) {
            ok = false;
            return;
        }
        a.all.push_back(*p);
        ok = hipMemset(*p, 0, sizeof(uint16_t) * cnt) == hipSuccess;
    };
    auto alf = [&](float** p, long cnt) {
        if (!ok || hipMalloc((void**)p, sizeof(float) * cnt) != hipSuccess) {
            ok = false;
            return;
        }
        a.all.push_back(*p);
    };
    al(&a.x0, 0, 2);  al(&a.e1a, 0, 8);  al(&a.e1, 0, 8);   al(&a.p1, 1, 8);
    al(&a.e2a, 1, 12); al(&a.e2, 1, 12); al(&a.p2, 2, 12);  al(&a.e3a, 2, 16);
    al(&a.e3, 2, 16); al(&a.p3, 3, 16);  al(&a.e4a, 3, 24); al(&a.e4, 3, 24);
    al(&a.r5, 2, 16); al(&a.d3, 2, 16);  al(&a.r6, 1, 12);  al(&a.d2, 1, 12);
    al(&a.r7, 0, 8);  al(&a.d1, 0, 8);
    al(&a.gout, 0, 1); al(&a.gd1, 0, 8); al(&a.gd2, 1, 12); al(&a.gd3, 2, 16);
    al(&a.ge1, 0, 8); al(&a.ge2, 1, 12); al(&a.ge3, 2, 16); al(&a.ge4, 3, 24);
    al(&a.ge1a, 0, 8); al(&a.ge2a, 1, 12); al(&a.ge3a, 2, 16); al(&a.ge4a, 3, 24);
    al(&a.gp1, 1, 8); al(&a.gp2, 2, 12); al(&a.gp3, 3, 16);
    al(&a.s0, 0, 24); al(&a.s1, 0, 24);
    alf(&a.out, vox(n));
    alf(&a.mono, vox(n));
    alf(&a.stats, 4);
    alf(&a.loss, 4);
    if (!ok) {
        fprintf(stderr, "hydro_mi355x: out of device memory for the network at %d^3\n", n);
        free_acts(a);
        return -2;
    }
    return 0;
}

int repack(Model& M) {
    for (int l = 0; l < kNL; ++l) {
        const LayerDef& L = kLayers[l];
        HYD_CHECK(pack_weights(M.params + M.woff[l], L.cin, L.cout, 0, M.wp[l], M.st));
        HYD_CHECK(pack_weights(M.params + M.woff[l], L.cin, L.cout, 1, M.wpt[l], M.st));
    }
    return 0;
}

int conv(Model& M, int l, int n_out, const uint16_t* x, const uint16_t* skip, uint16_t* yr,
         uint16_t* yd, float* yf, const float* add) {
    const LayerDef& L = kLayers[l];
    ConvArgs a;
    a.x = x;
    a.wp = M.wp[l];
    a.bias = M.params + M.boff[l];
    a.skip = skip;
    a.yr = yr;
    a.yd = yd;
    a.yf = yf;
    a.add = add;
    a.n = n_out;
    a.up = L.up;
    a.cinp = pad4(L.cin);
    a.cout = L.cout;
    a.coutp = pad4(L.cout);
    a.relu = L.relu;
    HYD_CHECK(conv_forward(a, M.st));
    return 0;
}

// Input gradient of layer l: the adjoint convolution of g (gradient with
// respect to the pre-activation, output grid) into out (output grid; for an
// upsampling layer this is the gradient of the upsampled input).
int conv_adj(Model& M, int l, int n_out, const uint16_t* g, uint16_t* out,
             const uint16_t* mask = nullptr) {
    const LayerDef& L = kLayers[l];
    ConvArgs a;
    a.x = g;
    a.wp = M.wpt[l];
    a.bias = nullptr;
    a.skip = nullptr;
    a.yr = out;
    a.yd = nullptr;
    a.yf = nullptr;
    a.add = nullptr;
    a.n = n_out;
    a.up = 0;
    a.cinp = pad4(L.cout);
    a.cout = L.cin;
    a.coutp = pad4(L.cin);
    a.relu = 0;
    a.mask = mask;
    HYD_CHECK(conv_forward(a, M.st));
    return 0;
}

int wgrad(Model& M, int l, int n_out, const uint16_t* x, const uint16_t* g) {
    const LayerDef& L = kLayers[l];
    WgradArgs a;
    a.x = x;
    a.g = g;
    a.dw = M.grads + M.woff[l];
    a.db = M.grads + M.boff[l];
    a.n = n_out;
    a.up = L.up;
    a.cinp = pad4(L.cin);
    a.cout = L.cout;
    a.coutp = pad4(L.cout);
    a.part = M.wg_part;
    HYD_CHECK(conv_wgrad(a, M.st));
    return 0;
}

// forward pass on the features in acts.x0 / acts.mono (grid edge n)
int forward(Model& M, int n) {
    Acts& A = M.acts;
    const int n1 = n / 2, n2 = n / 4, n3 = n / 8;
    if (conv(M, 0, n, A.x0, nullptr, A.e1a, nullptr, nullptr, nullptr) ||
        conv(M, 1, n, A.e1a, nullptr, A.e1, nullptr, nullptr, nullptr))
        return -2;
    HYD_CHECK(avgpool2(A.e1, n, 8, A.p1, M.st));
    if (conv(M, 2, n1, A.p1, nullptr, A.e2a, nullptr, nullptr, nullptr) ||
        conv(M, 3, n1, A.e2a, nullptr, A.e2, nullptr, nullptr, nullptr))
        return -2;
    HYD_CHECK(avgpool2(A.e2, n1, 12, A.p2, M.st));
    if (conv(M, 4, n2, A.p2, nullptr, A.e3a, nullptr, nullptr, nullptr) ||
        conv(M, 5, n2, A.e3a, nullptr, A.e3, nullptr, nullptr, nullptr))
        return -2;
    HYD_CHECK(avgpool2(A.e3, n2, 16, A.p3, M.st));
    if (conv(M, 6, n3, A.p3, nullptr, A.e4a, nullptr, nullptr, nullptr) ||
        conv(M, 7, n3, A.e4a, nullptr, A.e4, nullptr, nullptr, nullptr) ||
        conv(M, 8, n2, A.e4, A.e3, A.r5, A.d3, nullptr, nullptr) ||
        conv(M, 9, n1, A.d3, A.e2, A.r6, A.d2, nullptr, nullptr) ||
        conv(M, 10, n, A.d2, A.e1, A.r7, A.d1, nullptr, nullptr) ||
        conv(M, 11, n, A.d1, nullptr, nullptr, nullptr, A.out, A.mono))
        return -2;
    return 0;
}

// backward pass from acts.gout (gradient of the loss w.r.t. out); adds the
// parameter gradients into M.grads
int backward(Model& M, int n) {
    Acts& A = M.acts;
    const int n1 = n / 2, n2 = n / 4, n3 = n / 8;
    const long v0 = vox(n), v1 = vox(n1), v2 = vox(n2), v3 = vox(n3);
    // c8: d1 -> out (no ReLU)
    if (wgrad(M, 11, n, A.d1, A.gout) || conv_adj(M, 11, n, A.gout, A.gd1)) return -2;
    // c7 (up): d2 -> r7, d1 = r7 + e1
    HYD_CHECK(relu_mask(A.gd1, A.r7, v0 * 8, A.s0, M.st));
    if (wgrad(M, 10, n, A.d2, A.s0) || conv_adj(M, 10, n, A.s0, A.s1)) return -2;
    HYD_CHECK(sumpool2_add(A.s1, n, 12, A.gd2, 0, M.st));
    HYD_CHECK(copy_bf16(A.gd1, v0 * 8, A.ge1, M.st));
    // c6 (up): d3 -> r6, d2 = r6 + e2
    HYD_CHECK(relu_mask(A.gd2, A.r6, v1 * 12, A.s0, M.st));
    if (wgrad(M, 9, n1, A.d3, A.s0) || conv_adj(M, 9, n1, A.s0, A.s1)) return -2;
    HYD_CHECK(sumpool2_add(A.s1, n1, 16, A.gd3, 0, M.st));
    HYD_CHECK(copy_bf16(A.gd2, v1 * 12, A.ge2, M.st));
    // c5 (up): e4 -> r5, d3 = r5 + e3
    HYD_CHECK(relu_mask(A.gd3, A.r5, v2 * 16, A.s0, M.st));
    if (wgrad(M, 8, n2, A.e4, A.s0) || conv_adj(M, 8, n2, A.s0, A.s1)) return -2;
    HYD_CHECK(sumpool2_add(A.s1, n2, 24, A.ge4, 0, M.st));
    HYD_CHECK(copy_bf16(A.gd3, v2 * 16, A.ge3, M.st));
    // c4b, c4a
    HYD_CHECK(relu_mask(A.ge4, A.e4, v3 * 24, A.s0, M.st));
    // ge4a: the input gradient of c4b with c4a's ReLU derivative applied
    if (wgrad(M, 7, n3, A.e4a, A.s0) || conv_adj(M, 7, n3, A.s0, A.ge4a, A.e4a)) return -2;
    if (wgrad(M, 6, n3, A.p3, A.ge4a) || conv_adj(M, 6, n3, A.ge4a, A.gp3)) return -2;
    HYD_CHECK(unpool_add(A.gp3, n2, 16, A.ge3, M.st));
    // c3b, c3a
    HYD_CHECK(relu_mask(A.ge3, A.e3, v2 * 16, A.s0, M.st));
    if (wgrad(M, 5, n2, A.e3a, A.s0) || conv_adj(M, 5, n2, A.s0, A.ge3a, A.e3a)) return -2;
    if (wgrad(M, 4, n2, A.p2, A.ge3a) || conv_adj(M, 4, n2, A.ge3a, A.gp2)) return -2;
    HYD_CHECK(unpool_add(A.gp2, n1, 12, A.ge2, M.st));
    // c2b, c2a
    HYD_CHECK(relu_mask(A.ge2, A.e2, v1 * 12, A.s0, M.st));
    if (wgrad(M, 3, n1, A.e2a, A.s0) || conv_adj(M, 3, n1, A.s0, A.ge2a, A.e2a)) return -2;
    if (wgrad(M, 2, n1, A.p1, A.ge2a) || conv_adj(M, 2, n1, A.ge2a, A.gp1)) return -2;
    HYD_CHECK(unpool_add(A.gp1, n, 8, A.ge1, M.st));
    // c1b, c1a
    HYD_CHECK(relu_mask(A.ge1, A.e1, v0 * 8, A.s0, M.st));
    if (wgrad(M, 1, n, A.e1a, A.s0) || conv_adj(M, 1, n, A.s0, A.ge1a, A.e1a)) return -2;
    if (wgrad(M, 0, n, A.x0, A.ge1a)) return -2;
    return 0;
}

// features of a density on the box [x0, x0 + L]^3 with moments mom (device)
template <class T>
int features(Model& M, const T* rho, int n, const double* mom, const double box[4]) {
    HYD_CHECK(hipMemsetAsync(M.acts.stats, 0, sizeof(float) * 4, M.st));
    if (sizeof(T) == sizeof(float)) {
        HYD_CHECK(log_stats((const float*)rho, vox(n), M.acts.stats, M.st));
        HYD_CHECK(make_features((const float*)rho, n, M.acts.stats, mom, box[0], box[1], box[2],
                                box[3], M.acts.x0, M.acts.mono, M.st));
    } else {
        HYD_CHECK(log_stats_f64((const double*)rho, vox(n), M.acts.stats, M.st));
        HYD_CHECK(make_features_f64((const double*)rho, n, M.acts.stats, mom, box[0], box[1],
                                    box[2], box[3], M.acts.x0, M.acts.mono, M.st));
    }
    return 0;
}

// features of the start correction from an averaged residual r (n^3 FP32)
int features_corr(Model& M, const float* r, int n) {
    HYD_CHECK(corr_features(r, n, M.sumsq, M.sumsq_part, M.acts.x0, M.acts.mono, M.st));
    return 0;
}

bool grid_ok(int n) { return n >= 16 && n % 8 == 0; }

}  // namespace nnh

// Training set held on the device: FP32 density, the potential in units of
// G M / L (natural layout) and the moments (M, centre of mass) per field.
struct NNDataset {
    int n = 0, count = 0;  // for the start correction: rho = residual, phi = target
    double box[4] = {0.0, 0.0, 0.0, 1.0};  // x0, y0, z0, L
    float *rho = nullptr, *phi = nullptr;
    double* mom = nullptr;    // [count][4]
    double* rho64 = nullptr;  // staging for the solver
    double *blob = nullptr, *mode = nullptr;
};

namespace nnh {
// input features of dataset field idx for the model's kind
int field_features(Model& M, NNDataset* D, int idx) {
    const long nv = vox(D->n);
    if (M.kind) return features_corr(M, D->rho + nv * idx, D->n);
    return features(M, D->rho + nv * idx, D->n, D->mom + 4 * idx, D->box);
}
}  // namespace nnh

extern "C" {

void* nn_model_create(unsigned long long seed) {
    using namespace nnh;
    Model* M = new Model();
    M->st = hyd_stream();
    if (!M->st) {
        delete M;
        return nullptr;
    }
    long off = 0;
    for (int l = 0; l < kNL; ++l) {
        M->woff[l] = off;
        off += 27L * pad4(kLayers[l].cin) * kLayers[l].cout;
        M->boff[l] = off;
        off += kLayers[l].cout;
    }
    M->nparam = off;
    // He initialisation of the real input channels (padded channels stay
    // zero); the last layer starts small so the output starts near the
    // monopole term
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kNL; ++l) {
        const LayerDef& L = kLayers[l];
        const double sd = (l == kNL - 1 ? 0.05 : 1.0) * sqrt(2.0 / (27.0 * L.cin));
        std::normal_distribution<double> nd(0.0, sd);
        for (int t = 0; t < 27; ++t)
            for (int c = 0; c < L.cin; ++c)
                for (int o = 0; o < L.cout; ++o)
                    h[M->woff[l] + ((long)t * pad4(L.cin) + c) * L.cout + o] = (float)nd(rng);
    }
    bool ok = hipMalloc((void**)&M->params, sizeof(float) * off) == hipSuccess &&
              hipMalloc((void**)&M->grads, sizeof(float) * off) == hipSuccess &&
              hipMalloc((void**)&M->wg_part, sizeof(float) * wgrad_scratch_len()) == hipSuccess &&
              hipMalloc((void**)&M->m, sizeof(float) * off) == hipSuccess &&
              hipMalloc((void**)&M->v, sizeof(float) * off) == hipSuccess &&
              hipMalloc((void**)&M->mom_tmp, sizeof(double) * 4) == hipSuccess &&
              hipMalloc((void**)&M->sumsq, sizeof(double) * 4) == hipSuccess &&
              hipMalloc((void**)&M->sumsq_part, sizeof(double) * 256) == hipSuccess;
    for (int l = 0; ok && l < kNL; ++l) {
        const LayerDef& L = kLayers[l];
        ok = hipMalloc((void**)&M->wp[l], sizeof(uint16_t) * packed_len(pad4(L.cin), L.cout)) ==
                 hipSuccess &&
             hipMalloc((void**)&M->wpt[l], sizeof(uint16_t) * packed_len(pad4(L.cout), L.cin)) ==
                 hipSuccess;
    }
    if (!ok || hipMemcpy(M->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) ||
        hipMemset(M->m, 0, sizeof(float) * off) || hipMemset(M->v, 0, sizeof(float) * off) ||
        hipMemset(M->grads, 0, sizeof(float) * off) || repack(*M) ||
        hipStreamSynchronize(M->st)) {
        fprintf(stderr, "nn_model_create: device allocation failed\n");
        nn_model_destroy(M);
        return nullptr;
    }
    return M;
}

void nn_model_destroy(void* h) {
    using namespace nnh;
    Model* M = (Model*)h;
    if (!M) return;
    free_acts(M->acts);
    void* ps[] = {M->params, M->grads,  M->m,       M->v,      M->mom_tmp,
                  M->wg_part, M->sumsq, M->rcoarse, M->interp1, M->interp2, M->sumsq_part};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (double* p : M->coarse) (void)hipFree(p);
    for (int l = 0; l < kNL; ++l) {
        if (M->wp[l]) (void)hipFree(M->wp[l]);
        if (M->wpt[l]) (void)hipFree(M->wpt[l]);
    }
    delete M;
}

long nn_model_num_params(void* h) {
    (void)h;
    long n = 0;
    for (int l = 0; l < nnh::kNL; ++l)
        n += 27L * nnh::kLayers[l].cin * nnh::kLayers[l].cout + nnh::kLayers[l].cout;
    return n;
}

long nn_model_param_len(void* h) { return ((nnh::Model*)h)->nparam; }

long nn_model_param_offset(void* h, int layer, int bias) {
    nnh::Model* M = (nnh::Model*)h;
    if (layer < 0 || layer >= nnh::kNL) return -1;
    return bias ? M->boff[layer] : M->woff[layer];
}

int nn_model_save(void* h, const char* path) {
    nnh::Model* M = (nnh::Model*)h;
    std::vector<float> p(M->nparam);
    HYD_CHECK(hipMemcpy(p.data(), M->params, sizeof(float) * M->nparam, hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "nn_model_save: cannot open %s\n", path);
        return -1;
    }
    // HYDNN002: the potential guess; HYDNN003: a network of another kind,
    // its kind and C after the header
    const int hdr[2] = {nnh::kNL, (int)M->nparam};
    bool ok = fwrite(M->kind ? "HYDNN003" : "HYDNN002", 1, 8, f) == 8 &&
              fwrite(hdr, sizeof(int), 2, f) == 2;
    if (ok && M->kind)
        ok = fwrite(&M->kind, sizeof(int), 1, f) == 1 &&
             fwrite(&M->cscale, sizeof(double), 1, f) == 1;
    ok = ok && fwrite(p.data(), sizeof(float), M->nparam, f) == (size_t)M->nparam;
    fclose(f);
    return ok ? 0 : -1;
}

int nn_model_load(void* h, const char* path) {
    nnh::Model* M = (nnh::Model*)h;
    FILE* f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "nn_model_load: cannot open %s\n", path);
        return -1;
    }
    char magic[8];
    int hdr[2], kind = 0;
    double cscale = 1.0;
    std::vector<float> p(M->nparam);
    bool ok = fread(magic, 1, 8, f) == 8;
    const bool v3 = ok && memcmp(magic, "HYDNN003", 8) == 0;
    ok = ok && (v3 || memcmp(magic, "HYDNN002", 8) == 0) &&
         fread(hdr, sizeof(int), 2, f) == 2 && hdr[0] == nnh::kNL && hdr[1] == (int)M->nparam;
    if (ok && v3)
        ok = fread(&kind, sizeof(int), 1, f) == 1 && fread(&cscale, sizeof(double), 1, f) == 1 &&
             kind == 1 && cscale > 0.0 && std::isfinite(cscale);
    ok = ok && fread(p.data(), sizeof(float), M->nparam, f) == (size_t)M->nparam;
    fclose(f);
    if (!ok) {
        fprintf(stderr, "nn_model_load: %s is not a weights file of this network\n", path);
        return -1;
    }
    HYD_CHECK(hipMemcpy(M->params, p.data(), sizeof(float) * M->nparam, hipMemcpyHostToDevice));
    if (nnh::repack(*M)) return -2;
    HYD_CHECK(hipStreamSynchronize(M->st));
    M->kind = kind;
    M->cscale = cscale;
    return 0;
}

int nn_model_set_kind(void* h, int kind, double cscale) {
    nnh::Model* M = (nnh::Model*)h;
    if ((kind != 0 && kind != 1) || !(cscale > 0.0) || !std::isfinite(cscale)) {
        fprintf(stderr, "nn_model_set_kind: kind must be 0 or 1 and C positive\n");
        return -1;
    }
    M->kind = kind;
    M->cscale = kind ? cscale : 1.0;
    return 0;
}

int nn_model_kind(void* h, double* cscale) {
    nnh::Model* M = (nnh::Model*)h;
    if (cscale) *cscale = M->cscale;
    return M->kind;
}

int nn_model_get_params(void* h, float* out) {
    nnh::Model* M = (nnh::Model*)h;
    HYD_CHECK(hipMemcpy(out, M->params, sizeof(float) * M->nparam, hipMemcpyDeviceToHost));
    return 0;
}

int nn_model_set_params(void* h, const float* in) {
    nnh::Model* M = (nnh::Model*)h;
    HYD_CHECK(hipMemcpy(M->params, in, sizeof(float) * M->nparam, hipMemcpyHostToDevice));
    if (nnh::repack(*M)) return -2;
    HYD_CHECK(hipStreamSynchronize(M->st));
    return 0;
}

// One forward and backward pass on a host density (n^3 FP32) on the box
// [0, L]^3 with moments mom4 = (M, xcm, ycm, zcm) and a host target, for
// checking the kernels against a CPU reference.  Outputs (each may be NULL):
// the loss (mean squared error), the network output (n^3), the input
// features (n^3 x 2, BF16 values widened to FP32) and the parameter
// gradients (nn_model_param_len floats, packed as the parameters).
int nn_model_forward_backward(void* h, const float* rho, const float* target, int n,
                              const double* mom4, double L, float* loss, float* out,
                              float* feat, float* grads) {
    using namespace nnh;
    Model* M = (Model*)h;
    if (!grid_ok(n)) {
        fprintf(stderr, "nn_model_forward_backward: n must be a multiple of 8, >= 16\n");
        return -1;
    }
    if (alloc_acts(M->acts, n)) return -2;
    const long nv = vox(n);
    float *drho = nullptr, *dtg = nullptr;
    HYD_CHECK(hipMalloc((void**)&drho, sizeof(float) * nv));
    HYD_CHECK(hipMalloc((void**)&dtg, sizeof(float) * nv));
    const double box[4] = {0.0, 0.0, 0.0, L};
    int rc = 0;
    if (hipMemcpy(drho, rho, sizeof(float) * nv, hipMemcpyHostToDevice) ||
        hipMemcpy(dtg, target, sizeof(float) * nv, hipMemcpyHostToDevice) ||
        hipMemcpy(M->mom_tmp, mom4, sizeof(double) * 4, hipMemcpyHostToDevice) ||
        hipMemsetAsync(M->grads, 0, sizeof(float) * M->nparam, M->st) ||
        hipMemsetAsync(M->acts.loss, 0, sizeof(float), M->st))
        rc = -2;
    if (!rc)
        rc = (M->kind ? features_corr(*M, drho, n) : features(*M, drho, n, M->mom_tmp, box)) ||
                     forward(*M, n)
                 ? -2
                 : 0;
    if (!rc && mse_loss(M->acts.out, dtg, nv, 1.0f / nv, M->acts.gout, M->acts.loss, M->st))
        rc = -2;
    if (!rc && backward(*M, n)) rc = -2;
    if (!rc && hipStreamSynchronize(M->st)) rc = -2;
    if (!rc && loss && hipMemcpy(loss, M->acts.loss, sizeof(float), hipMemcpyDeviceToHost))
        rc = -2;
    if (!rc && out && hipMemcpy(out, M->acts.out, sizeof(float) * nv, hipMemcpyDeviceToHost))
        rc = -2;
    if (!rc && grads &&
        hipMemcpy(grads, M->grads, sizeof(float) * M->nparam, hipMemcpyDeviceToHost))
        rc = -2;
    if (!rc && feat) {
        std::vector<uint16_t> x(nv * 4);
        if (hipMemcpy(x.data(), M->acts.x0, sizeof(uint16_t) * nv * 4, hipMemcpyDeviceToHost)) {
            rc = -2;
        } else {
            for (long v = 0; v < nv; ++v)
                for (int c = 0; c < 2; ++c) {
                    const uint32_t u = (uint32_t)x[4 * v + c] << 16;
                    memcpy(&feat[2 * v + c], &u, sizeof(float));
                }
        }
    }
    (void)hipFree(drho);
    (void)hipFree(dtg);
    return rc;
}

// ---- training data ------------------------------------------------------
void* nn_dataset_create(int n, int count) {
    if (!nnh::grid_ok(n) || count < 1) {
        fprintf(stderr, "nn_dataset_create: n must be a multiple of 8, >= 16\n");
        return nullptr;
    }
    NNDataset* D = new NNDataset();
    D->n = n;
    D->count = count;
    const long nv = nnh::vox(n);
    if (hipMalloc((void**)&D->rho, sizeof(float) * nv * count) != hipSuccess ||
        hipMalloc((void**)&D->phi, sizeof(float) * nv * count) != hipSuccess ||
        hipMalloc((void**)&D->mom, sizeof(double) * 4 * count) != hipSuccess ||
        hipMalloc((void**)&D->rho64, sizeof(double) * nv) != hipSuccess ||
        hipMalloc((void**)&D->blob, sizeof(double) * 6 * NN_MAX_BLOBS) != hipSuccess ||
        hipMalloc((void**)&D->mode, sizeof(double) * 5 * NN_MAX_MODES) != hipSuccess) {
        fprintf(stderr, "nn_dataset_create: out of device memory for %d fields of %d^3\n", count,
                n);
        nn_dataset_destroy(D);
        return nullptr;
    }
    return D;
}

void nn_dataset_destroy(void* h) {
    NNDataset* D = (NNDataset*)h;
    if (!D) return;
    void* ps[] = {D->rho, D->phi, D->mom, D->rho64, D->blob, D->mode};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete D;
}

// Generate field `index` (blob_field_kernel), solve it cold with `grav` (an
// n^3 solver: its omega, tolerance, criterion and Poisson mode apply) and
// store density, potential in units of G M / L, and moments.  Returns the
// solver status (0 converged).
int nn_dataset_fill(void* h, int index, const double* blob, int nblob, const double* mode,
                    int nmode, double floor_rho, double sigma, GravitySolver3D* grav) {
    NNDataset* D = (NNDataset*)h;
    const long nv = nnh::vox(D->n);
    if (index < 0 || index >= D->count || nblob < 0 || nblob > NN_MAX_BLOBS || nmode < 0 ||
        nmode > NN_MAX_MODES || grav->nx != D->n || grav->ny != D->n || grav->nz != D->n ||
        grav->bc_type != GRAV_BC_ISOLATED) {
        fprintf(stderr, "nn_dataset_fill: bad arguments\n");
        return -1;
    }
    GravDevice* g = grav->dev;
    hipStream_t st = g->stream;
    D->box[0] = grav->xmin;
    D->box[1] = grav->ymin;
    D->box[2] = grav->zmin;
    D->box[3] = grav->xmax - grav->xmin;
    if (nblob)
        HYD_CHECK(hipMemcpyAsync(D->blob, blob, sizeof(double) * 6 * nblob,
                                 hipMemcpyHostToDevice, st));
    if (nmode)
        HYD_CHECK(hipMemcpyAsync(D->mode, mode, sizeof(double) * 5 * nmode,
                                 hipMemcpyHostToDevice, st));
    HYD_CHECK(hyd::nn::blob_field(D->n, D->blob, nblob, D->mode, nmode, floor_rho, sigma, D->rho64,
                                  D->rho + nv * index, st));
    HYD_CHECK(hipMemsetAsync(g->phi_r, 0, sizeof(double) * g->nsplit, st));
    HYD_CHECK(hipMemsetAsync(g->phi_b, 0, sizeof(double) * g->nsplit, st));
    const int status = gravity_solve_device(grav, D->rho64);
    if (status < -1) return status;
    if (!g->nat_tmp && hipMalloc((void**)&g->nat_tmp, sizeof(double) * nv) != hipSuccess)
        return -2;
    HYD_CHECK(hyd::launch_split_to_natural(g->g, g->phi_r, g->phi_b, g->nat_tmp, st));
    HYD_CHECK(hyd::nn::scale_potential(g->nat_tmp, nv, g->mom, grav->G, D->box[3],
                                       D->phi + nv * index, st));
    HYD_CHECK(hipMemcpyAsync(D->mom + 4 * index, g->mom, sizeof(double) * 4,
                             hipMemcpyDeviceToDevice, st));
    HYD_CHECK(hipStreamSynchronize(st));
    return status;
}

// Field `index` from host arrays (n^3 FP32 each): the network input and its
// target.  For the start correction the input is the averaged residual and
// the target the correction in units of C L^2 s.
int nn_dataset_set(void* h, int index, const float* input, const float* target) {
    NNDataset* D = (NNDataset*)h;
    const long nv = nnh::vox(D->n);
    if (index < 0 || index >= D->count || !input || !target) {
        fprintf(stderr, "nn_dataset_set: bad arguments\n");
        return -1;
    }
    const double mom4[4] = {0.0, 0.0, 0.0, 0.0};
    HYD_CHECK(hipMemcpy(D->rho + nv * index, input, sizeof(float) * nv, hipMemcpyHostToDevice));
    HYD_CHECK(hipMemcpy(D->phi + nv * index, target, sizeof(float) * nv, hipMemcpyHostToDevice));
    HYD_CHECK(hipMemcpy(D->mom + 4 * index, mom4, sizeof(mom4), hipMemcpyHostToDevice));
    return 0;
}

int nn_dataset_get(void* h, int index, float* rho, float* phi, double* mom4) {
    NNDataset* D = (NNDataset*)h;
    const long nv = nnh::vox(D->n);
    if (index < 0 || index >= D->count) return -1;
    if (rho)
        HYD_CHECK(hipMemcpy(rho, D->rho + nv * index, sizeof(float) * nv, hipMemcpyDeviceToHost));
    if (phi)
        HYD_CHECK(hipMemcpy(phi, D->phi + nv * index, sizeof(float) * nv, hipMemcpyDeviceToHost));
    if (mom4)
        HYD_CHECK(hipMemcpy(mom4, D->mom + 4 * index, sizeof(double) * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Adam steps over mini-batches: order[] lists field indices (steps * batch
// entries); the gradient of the batch-mean loss, then one Adam update
// (beta1 0.9, beta2 0.999, eps 1e-8).  loss_log (may be NULL) receives the
// batch-mean loss of every step.
int nn_train(void* hm, void* hd, const int* order, int steps, int batch, float lr,
             float* loss_log) {
    using namespace nnh;
    Model* M = (Model*)hm;
    NNDataset* D = (NNDataset*)hd;
    const int n = D->n;
    const long nv = vox(n);
    if (steps < 0 || batch < 1) return -1;
    for (long i = 0; i < (long)steps * batch; ++i)
        if (order[i] < 0 || order[i] >= D->count) return -1;
    if (alloc_acts(M->acts, n)) return -2;
    for (int s = 0; s < steps; ++s) {
        HYD_CHECK(hipMemsetAsync(M->grads, 0, sizeof(float) * M->nparam, M->st));
        HYD_CHECK(hipMemsetAsync(M->acts.loss, 0, sizeof(float), M->st));
        for (int b = 0; b < batch; ++b) {
            const int idx = order[s * batch + b];
            if (field_features(*M, D, idx) || forward(*M, n)) return -2;
            HYD_CHECK(mse_loss(M->acts.out, D->phi + nv * idx, nv, 1.0f / ((float)nv * batch),
                               M->acts.gout, M->acts.loss, M->st));
            if (backward(*M, n)) return -2;
        }
        M->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - powf(b1, (float)M->step), c2 = 1.0f - powf(b2, (float)M->step);
        HYD_CHECK(adam_step(M->params, M->grads, M->m, M->v, M->nparam, lr, b1, b2, 1e-8f, c1, c2,
                            M->st));
        if (repack(*M)) return -2;
        if (loss_log) {
            HYD_CHECK(hipMemcpyAsync(&loss_log[s], M->acts.loss, sizeof(float),
                                     hipMemcpyDeviceToHost, M->st));
        }
    }
    HYD_CHECK(hipStreamSynchronize(M->st));
    return 0;
}

// mean squared error of the network on dataset fields [first, first + count)
int nn_evaluate(void* hm, void* hd, int first, int count, float* mse) {
    using namespace nnh;
    Model* M = (Model*)hm;
    NNDataset* D = (NNDataset*)hd;
    const int n = D->n;
    const long nv = vox(n);
    if (first < 0 || count < 1 || first + count > D->count) return -1;
    if (alloc_acts(M->acts, n)) return -2;
    HYD_CHECK(hipMemsetAsync(M->acts.loss, 0, sizeof(float), M->st));
    for (int i = first; i < first + count; ++i) {
        if (field_features(*M, D, i) || forward(*M, n)) return -2;
        HYD_CHECK(mse_loss(M->acts.out, D->phi + nv * i, nv, 1.0f / ((float)nv * count),
                           M->acts.gout, M->acts.loss, M->st));
    }
    HYD_CHECK(hipMemcpyAsync(mse, M->acts.loss, sizeof(float), hipMemcpyDeviceToHost, M->st));
    HYD_CHECK(hipStreamSynchronize(M->st));
    return 0;
}

// Device time of the network on an n^3 grid (features of a synthetic
// density already on the device): the average of `reps` forward passes and
// of `reps` forward + backward passes (loss against a zero target), in ms.
int nn_benchmark(void* hm, int n, int reps, double* fwd_ms, double* fwdbwd_ms) {
    using namespace nnh;
    Model* M = (Model*)hm;
    if (!grid_ok(n) || reps < 1) return -1;
    if (alloc_acts(M->acts, n)) return -2;
    const long nv = vox(n);
    float *rho = nullptr, *tgt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc((void**)&rho, sizeof(float) * nv) || hipMalloc((void**)&tgt, sizeof(float) * nv) ||
        hipEventCreate(&e0) || hipEventCreate(&e1))
        rc = -2;
    // a smooth positive density and the moments of a unit mass at the centre
    const double mom[4] = {1.0, 0.5, 0.5, 0.5};
    const double box[4] = {0.0, 0.0, 0.0, 1.0};
    std::vector<float> h(nv);
    for (long v = 0; v < nv; ++v) h[v] = 1.0f + 0.5f * (float)((v * 2654435761u) % 1000) / 1000.0f;
    if (!rc && (hipMemcpy(rho, h.data(), sizeof(float) * nv, hipMemcpyHostToDevice) ||
                hipMemset(tgt, 0, sizeof(float) * nv) ||
                hipMemcpy(M->mom_tmp, mom, sizeof(mom), hipMemcpyHostToDevice)))
        rc = -2;
    if (!rc && (features(*M, rho, n, M->mom_tmp, box) || forward(*M, n))) rc = -2;  // warm-up
    float ms = 0.f;
    if (!rc && hipEventRecord(e0, M->st)) rc = -2;
    for (int r = 0; !rc && r < reps; ++r)
        if (forward(*M, n)) rc = -2;
    if (!rc && (hipEventRecord(e1, M->st) || hipEventSynchronize(e1) ||
                hipEventElapsedTime(&ms, e0, e1)))
        rc = -2;
    if (!rc) *fwd_ms = ms / reps;
    if (!rc && hipEventRecord(e0, M->st)) rc = -2;
    for (int r = 0; !rc && r < reps; ++r) {
        if (forward(*M, n) || mse_loss(M->acts.out, tgt, nv, 1.0f / nv, M->acts.gout,
                                       M->acts.loss, M->st) ||
            backward(*M, n))
            rc = -2;
    }
    if (!rc && (hipEventRecord(e1, M->st) || hipEventSynchronize(e1) ||
                hipEventElapsedTime(&ms, e0, e1)))
        rc = -2;
    if (!rc) *fwdbwd_ms = ms / reps;
    if (rho) (void)hipFree(rho);
    if (tgt) (void)hipFree(tgt);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

// Multiply-add flops of one forward pass on an n^3 grid counting real
// channels only: 2 * 27 * cin * cout per output voxel of each convolution.
double nn_forward_flops(int n) {
    using namespace nnh;
    const int lvl[kNL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0, 0};  // output grid level
    double f = 0.0;
    for (int l = 0; l < kNL; ++l)
        f += 2.0 * 27.0 * kLayers[l].cin * kLayers[l].cout * (double)vox(n >> lvl[l]);
    return f;
}

}  // extern "C"

// Potential guess for a device density on a cubic n^3 grid with the box
// [x0, x0 + L]^3 and moments mom (device, M and centre of mass first): the
// density is coarsened by 2 until the grid is at most 128^3, the network
// runs there, and its output is interpolated back to n^3 and scaled by
// G M / L into phi_nat (natural layout, FP64).  Returns -1 when the grid
// does not fit the network (no guess written).
int nn_guess_device(void* hm, const double* d_rho, int n, const double* d_mom, double G,
                    const double box[4], double* d_phi_nat) {
    using namespace nnh;
    Model* M = (Model*)hm;
    int nc = n;
    while (nc > 128 && nc % 2 == 0) nc /= 2;
    if (nc > 128 || !grid_ok(nc)) return -1;
    const double* cur = d_rho;
    int cn = n, level = 0;
    while (cn > nc) {
        if ((int)M->coarse.size() <= level) {
            double* p = nullptr;
            HYD_CHECK(hipMalloc((void**)&p, sizeof(double) * vox(cn / 2)));
            M->coarse.push_back(p);
            M->coarse_n.push_back(cn / 2);
        } else if (M->coarse_n[level] != cn / 2) {
            (void)hipFree(M->coarse[level]);
            HYD_CHECK(hipMalloc((void**)&M->coarse[level], sizeof(double) * vox(cn / 2)));
            M->coarse_n[level] = cn / 2;
        }
        HYD_CHECK(restrict2_f64(cur, cn, M->coarse[level], M->st));
        cur = M->coarse[level];
        cn /= 2;
        ++level;
    }
    if (alloc_acts(M->acts, nc)) return -2;
    if (features(*M, cur, nc, d_mom, box) || forward(*M, nc)) return -2;
    HYD_CHECK(prolong_to_f64(M->acts.out, nc, n, d_mom, G, box[3], d_phi_nat, M->st));
    return 0;
}

// Correction of the start of a cubic n^3 solve with box edge L: the split
// FP32 residual (rr, rb) of the start is averaged onto nc^3 (nc = n halved
// until at most 64), the start-correction network runs there, and
// e = C L^2 s P(out) on the interior, 0 on the boundary (split layout,
// FP32), P cubic (Catmull-Rom) interpolation.  Returns -1 when the grid does
// not fit the network.
int nn_correction_device(void* hm, const float* rr, const float* rb, int n, double L, float* er,
                         float* eb) {
    using namespace nnh;
    Model* M = (Model*)hm;
    if (M->kind != 1 || n % 2 != 0) return -1;
    int nc = n;
    while (nc > 64 && nc % 2 == 0) nc /= 2;
    if (nc > 64 || !grid_ok(nc)) return -1;
    if (M->rcoarse_n != nc) {
        if (M->rcoarse) (void)hipFree(M->rcoarse);
        M->rcoarse = nullptr;
        M->rcoarse_n = 0;
        HYD_CHECK(hipMalloc((void**)&M->rcoarse, sizeof(float) * vox(nc)));
        M->rcoarse_n = nc;
    }
    if (M->interp_n != n) {
        if (M->interp1) (void)hipFree(M->interp1);
        if (M->interp2) (void)hipFree(M->interp2);
        M->interp1 = M->interp2 = nullptr;
        M->interp_n = 0;
        HYD_CHECK(hipMalloc((void**)&M->interp1, sizeof(float) * nc * nc * n));
        HYD_CHECK(hipMalloc((void**)&M->interp2, sizeof(float) * nc * (long)n * n));
        M->interp_n = n;
    }
    if (alloc_acts(M->acts, nc)) return -2;
    HYD_CHECK(corr_restrict(rr, rb, n, n / nc, M->rcoarse, M->st));
    if (features_corr(*M, M->rcoarse, nc) || forward(*M, nc)) return -2;
    HYD_CHECK(corr_apply(M->acts.out, nc, n, M->interp1, M->interp2, M->sumsq,
                         M->cscale * L * L, er, eb, M->st));
    return 0;
}
