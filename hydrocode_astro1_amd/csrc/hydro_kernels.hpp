// hydro_kernels.hpp -- launch interface of the hydro step kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace hyd {

struct HydroStepArgs {
    const double* uin[5];   // state planes being advanced
    double* uout[5];        // advanced state (interior cells written here)
    const double* g[3];     // gravitational acceleration planes (gravity != 0)
    double* face;           // boundary cells of U_pre (see face_index)
    // Signal-speed slots (ordered bits) and their NaN flags: 0 no NaN speed,
    // 1 a NaN speed was seen, 2 the slot holds the reference's fold over the
    // speeds after the last NaN (nan_fold_block).
    const unsigned long long* ms_in;  // signal speed of uin
    unsigned long long* ms_out;       // signal speed of uout, accumulated
    unsigned long long* ms_clear;     // slot reset for the step after next
    const unsigned long long* nan_in;
    unsigned long long* nan_out;
    unsigned long long* nan_clear;
    double* dt_out;         // dt used by this step
    int nx, ny, nz;
    int cx;                 // (tile, plane) pairs per block; 0: the launcher chooses
    int ntj, ntk, wpb;      // filled by the launcher (wpb: pairs per block)
    double dx, dy, dz, cfl, gamma;
    int gravity;
    int zcopy;              // z faces outflow on both sides: the step writes them (hydro_zcopy)
};

struct HydroBCArgs {
    double* u[5];           // state being finalised in place
    const double* face;     // U_pre of the boundary cells
    unsigned long long* ms_out;
    unsigned long long* nan_out;  // NaN flag of ms_out
    int nx, ny, nz;
    int bc[6];              // xmin xmax ymin ymax zmin zmax (BCType values)
    double gamma;
    int zcopy;              // the step kernel has written the z faces: x and y faces only
};

// True when both z sides are outflow and each z boundary cell shares a
// wavefront with its inward neighbour: the step kernel then writes the z
// faces' final values itself (a lane shuffle of the neighbour's new state)
// and the boundary kernel leaves them alone.
bool hydro_zcopy(int nz, int bc_zmin, int bc_zmax);

// face buffer size in doubles for an nx x ny x nz grid
inline long face_buffer_len(int nx, int ny, int nz) {
    return 10L * ny * nz + 10L * nx * nz + 10L * nx * ny;
}

hipError_t launch_hydro_step(const HydroStepArgs& a, hipStream_t st);
hipError_t launch_hydro_bc(const HydroBCArgs& a, hipStream_t st);
// Signal speed of a whole state into *out with its NaN flag (both zero on
// entry), the reference's order kept when a speed is NaN.
hipError_t launch_hydro_maxspeed(double* const u[5], long n, double gamma,
                                 unsigned long long* out, unsigned long long* nan_flag,
                                 hipStream_t st);
hipError_t launch_aos_to_soa(const double* aos, double* const u[5], long n, hipStream_t st);
hipError_t launch_soa_to_aos(double* aos, double* const u[5], long n, hipStream_t st);

}  // namespace hyd
