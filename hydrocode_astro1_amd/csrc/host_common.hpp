// host_common.hpp -- host-side helpers shared by the C-ABI implementation:
// HIP error mapping (report on stderr, return an error code, never abort)
// and the device contexts hung off the public structs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>

#include <vector>

#include "gravity_kernels.hpp"
#include "hydro_kernels.hpp"
#include "multigrid.hpp"
#include "seqsum.hpp"

#define HYD_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "hydro_mi355x: %s failed at %s:%d: %s\n", #expr, __FILE__,   \
                    __LINE__, hipGetErrorString(e_));                                    \
            return -2;                                                                   \
        }                                                                                \
    } while (0)

namespace hyd {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace hyd

// Device context of a GravitySolver3D.
struct GravDevice {
    hyd::GridDims g;
    long n = 0;       // cells
    long nsplit = 0;  // entries per colour array
    double *phi_r = nullptr, *phi_b = nullptr;
    double *rhs_r = nullptr, *rhs_b = nullptr;
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;
    double *rho_tmp = nullptr;  // device copy of a host density
    double *nat_tmp = nullptr;  // natural-layout staging
    double *mom = nullptr;      // moment results
    double *mom_scratch = nullptr;
    long mom_scratch_len = 0;
    double *mom_seq = nullptr;  // stored terms and sums of the sequential order
    long mom_seq_len = 0;
    int seq_method = hyd::SEQSUM_PEAKS;  // HYDRO_MI355X_SEQSUM=chain or checked
    double *wlm = nullptr;      // multipole weights
    hyd::SorCtl *ctl = nullptr;
    hyd::SorCtl *ctl_host = nullptr;  // pinned readback
    double *acc = nullptr;            // residual sums (device)
    double *res_part = nullptr;       // per-block residual sums
    double *acc_host = nullptr;       // pinned readback
    float *e_r = nullptr, *e_b = nullptr, *r_r = nullptr, *r_b = nullptr;
    double *guess_r = nullptr, *guess_b = nullptr;  // learned guess (split layout)
    bool owns_nn = false;  // the network was created from an input deck
    bool owns_corr = false;  // the same for the start-correction network
    // set by the start choice of POISSON_NN_FP32_DC when r_r / r_b already
    // hold the chosen start's residual (boundary applied) and start_rel its
    // relative residual; the defect correction then skips recomputing them
    bool start_residual_ready = false;
    double start_rel = 0.0;
    // relative residual of the potential guess the last time it was tried
    // (-1: never); the guess is skipped while the best start so far is under
    // a quarter of it
    double guess_last_rel = -1.0;
    // multigrid inner solve of the defect correction (DC_INNER_MULTIGRID):
    // pre- and post-smoothing sweeps per level and their SOR factor; V(1,1)
    // with 1.15 reached 1e-8 fastest at 256^3 (5.9 ms against 7.7 for
    // V(2,2) with 1.0, 6.1 for V(1,2) and 6.5 for V(2,1) with 1.15)
    int mg_nu1 = 1, mg_nu2 = 1;
    float mg_omega = 1.15f;
    std::vector<hyd::MgLevel> mg;          // levels 1, 2, ... (coarser along the list)
    std::vector<float*> mg_u, mg_f, mg_q;  // correction, right-hand side, residual
    float *mg_qr = nullptr, *mg_qb = nullptr;  // level-0 residual (split layout)
    // levels from mg_tail_first down run in one workgroup in LDS (-1: none);
    // HYDRO_MI355X_MG_TAIL=0 runs them one phase per launch
    bool mg_use_tail = true;
    int mg_tail_first = -1, mg_tail_floats = 0;
    hyd::MgTail* mg_tail_dev = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool g_valid = false;  // device g has been computed at least once
    // nz == 1 (gravity_solver_3d.c:569-634): phi_r holds the potential in
    // the natural layout; the split arrays and the defect-correction modes
    // are not used
    bool two_d = false;
    const double* rho_2d = nullptr;  // density of the solve under way (nz == 1)
};

struct OutputPipe;  // hydro_io.hip: snapshot, copy stream and writer thread

// Device context of a GodunovHLLC3D.
struct HydroDevice {
    int nx = 0, ny = 0, nz = 0;
    long n = 0;
    double *U[2][5] = {{nullptr}};
    int cur = 0;
    double *face = nullptr;
    // [0..2] signal-speed slots, [3..5] their NaN flags, [6] unused
    unsigned long long *ms = nullptr;
    double *dt_log = nullptr;          // [DT_LOG] dt per step (ring)
    double *g_zero = nullptr;          // zero g planes when gravity is off
    double *aos = nullptr;             // staging for AoS transfers
    long step = 0;                     // steps since the last upload
    int cx = 0;                        // x planes per step-kernel block (0: chosen per grid)
    bool uploaded = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_hydro_ms = 0.0;
    hipEvent_t record_ev[3] = {nullptr, nullptr, nullptr};  // benchmark brackets
    bool skip_gravity = false;  // gravity already solved by the caller
    OutputPipe *out = nullptr;  // asynchronous output of the run loops
    static constexpr int DT_LOG = 1 << 16;
    static constexpr int MS_WORDS = 8;
};

// Learned potential guess of a cubic device density (nn_api.hip); -1 when
// the grid does not fit the network.  box = x0, y0, z0, L.
int nn_guess_device(void *model, const double *d_rho, int n, const double *d_mom, double G,
                    const double box[4], double *d_phi_nat);
// The learned correction of a start from its split FP32 residual (rr, rb)
// into (er, eb), split FP32; -1 when the grid does not fit the network.
int nn_correction_device(void *model, const float *rr, const float *rb, int n, double L,
                         float *er, float *eb);

// Internal interface between the hydro driver and the gravity solver.
struct GravitySolver3D;
int grav_internal_solve(struct GravitySolver3D *s, const double *d_rho, hipStream_t st);
const double *grav_internal_g(struct GravitySolver3D *s, int comp);
// Device potential (natural order) and g from host arrays; NULL leaves that
// part as it is.  The next solve warm-starts from the potential, and g is
// what a step uses when its solve does not converge.
int grav_internal_restore(struct GravitySolver3D *s, const double *phi, const double *gx,
                          const double *gy, const double *gz);

// Between the hydro driver (hydro_api.hip) and its files (hydro_io.hip).
struct GodunovHLLC3D;
int hyd_download(struct GodunovHLLC3D *s);  // device state (and g) to the host mirror
int hyd_output_async(struct GodunovHLLC3D *s, int count);
void hyd_output_drain(HydroDevice *d);  // wait for the pending output files
void hyd_output_free(HydroDevice *d);
// HYDIC001 / HYDCK001 file into s; t and count from a checkpoint (else 0).
int hyd_read_state_file(struct GodunovHLLC3D *s, const char *path, double *t, int *count);
