// noalloc.hip -- the multipole moments of a device density with no device
// allocation: the same kernels and summation tree as gravity_moments (tree
// order), with their block partials in a file-scope device array instead of
// a buffer the library allocates.  For callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The array `scratch` exists once per GPU the library runs on and is held
// until the process ends.  Host code never passes its name to a kernel (on
// the host that name is only a stand-in): the device address is looked up
// once per device with hipGetSymbolAddress and kept in scratch_ptr[].  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// (from several host threads) race; each call finishes its work before it
// returns.
#include <stdio.h>

#include <mutex>

#include "../../include/hydro_mi355x.h"
#include "host_common.hpp"

hipStream_t hyd_stream();

namespace {

constexpr long kScratchDoubles = GRAVITY_MOMENTS_NOALLOC_MAX_SCRATCH;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

double* scratch_ptr[kMaxDevices] = {nullptr};
std::mutex scratch_mutex;

// The device address of this GPU's scratch array (looked up on the first
// call on each device, with that device current), or nullptr.
double* device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fprintf(stderr, "gravity_moments_noalloc: no usable current device\n");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "gravity_moments_noalloc: hipGetSymbolAddress failed: %s\n",
                    hipGetErrorString(e));
            return nullptr;
        }
        scratch_ptr[dev] = static_cast<double*>(p);
    }
    return scratch_ptr[dev];
}

}  // namespace

extern "C" {

long gravity_moments_noalloc_scratch(int nx, int ny, int nz, int lmax) {
    if (nx < 1 || ny < 1 || nz < 1 || lmax < 2 || lmax > hyd::MP_LMAX_MAX) return -1;
    return hyd::moments_scratch_len((long)nx * ny * nz, lmax);
}

int gravity_moments_noalloc(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                            double ymax, double zmin, double zmax, int lmax, const double* d_rho,
                            double* d_mom) {
    if (!d_rho || !d_mom || nx < 1 || ny < 1 || nz < 1) {
        fprintf(stderr, "gravity_moments_noalloc: bad arguments\n");
        return -1;
    }
    if (lmax < 2 || lmax > hyd::MP_LMAX_MAX) {
        fprintf(stderr, "gravity_moments_noalloc: lmax %d outside 2..%d\n", lmax,
                hyd::MP_LMAX_MAX);
        return -1;
    }
    const long need = hyd::moments_scratch_len((long)nx * ny * nz, lmax);
    if (need > kScratchDoubles) {
        fprintf(stderr,
                "gravity_moments_noalloc: a %d x %d x %d grid at lmax %d needs %ld scratch "
                "doubles, more than the %ld this entry point holds\n",
                nx, ny, nz, lmax, need, kScratchDoubles);
        return -1;
    }
    double* part = device_scratch();
    hipStream_t st = hyd_stream();
    if (!part || !st) return -2;
    // the grid as gravity_solver_create sets it up (gravity_solver_3d.c:60-86)
    hyd::GridDims g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.nzh = (nz + 1) / 2;
    g.xmin = xmin; g.xmax = xmax; g.ymin = ymin; g.ymax = ymax; g.zmin = zmin; g.zmax = zmax;
    g.dx = (xmax - xmin) / nx;
    g.dy = (ymax - ymin) / ny;
    g.dz = (zmax - zmin) / nz;
    HYD_CHECK(hyd::launch_moments(g, d_rho, lmax, part, d_mom, st));
    HYD_CHECK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
