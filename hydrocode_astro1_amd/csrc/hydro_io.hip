// hydro_io.hip -- files at the boundary of the hydro path.
//
//  - the reference's CSV and VTK writers (hydro_solver_3d.c:2336-2435), in
//    its text formats, from the host mirror or from a device snapshot;
//  - its CSV restart reader (hydro_solver_3d.c:2139-2235), restated below
//    from what a restart file means to it;
//  - this build's binary files: the initial state (HYDIC001, U only) and the
//    checkpoint (HYDCK001: U, the potential and g at full precision, so a run
//    resumed from it continues bit for bit, warm SOR start included);
//  - the output pipe of the run loops: the state is snapshot on the device,
//    copied to pinned host memory on a second stream while the steps go on,
//    and written by a host thread.
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "../../include/hydro_mi355x.h"
#include "host_common.hpp"

namespace {

// ---------------------------------------------------------------- EOS ----
// cons_to_prim_3d, hydro_solver_3d.c:162-175 (what the writers print)
Primitive3D c2p(const State3D& U, double gamma) {
    Primitive3D P;
    P.rho = (U.rho > RHO_FLOOR) ? U.rho : RHO_FLOOR;
    P.u = U.rhou / P.rho;
    P.v = U.rhov / P.rho;
    P.w = U.rhow / P.rho;
    double ke = 0.5 * P.rho * (P.u * P.u + P.v * P.v + P.w * P.w);
    P.p = (gamma - 1.0) * (U.E - ke);
    P.p = (P.p > P_FLOOR) ? P.p : P_FLOOR;
    return P;
}
// prim_to_cons_3d, hydro_solver_3d.c:177-189 (what the restart reader stores)
State3D p2c(const Primitive3D& P, double gamma) {
    State3D U;
    U.rho = P.rho;
    U.rhou = P.rho * P.u;
    U.rhov = P.rho * P.v;
    U.rhow = P.rho * P.w;
    double ke = 0.5 * P.rho * (P.u * P.u + P.v * P.v + P.w * P.w);
    U.E = P.p / (gamma - 1.0) + ke;
    return U;
}

// ------------------------------------------------------------ writers ----
struct Geometry {
    int nx, ny, nz;
    double xmin, ymin, zmin, dx, dy, dz, gamma;
    bool gravity;
};

Geometry geometry(const GodunovHLLC3D* s) {
    return Geometry{s->nx,   s->ny,   s->nz,   s->xmin,  s->ymin,          s->zmin,
                    s->dx,   s->dy,   s->dz,   s->gamma, s->enable_gravity};
}

// The fields the writers print, in either host layout: the reference's AoS
// mirror, or the SoA planes of a device snapshot.
struct HostFields {
    const State3D* aos = nullptr;
    const double* plane[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const double* g[3] = {nullptr, nullptr, nullptr};  // NULL: zero

    State3D cell(long c) const {
        if (aos) return aos[c];
        return State3D{plane[0][c], plane[1][c], plane[2][c], plane[3][c], plane[4][c]};
    }
    double accel(int d, long c) const { return g[d] ? g[d][c] : 0.0; }
};

FILE* open_output(int step, const char* ext) {
    char name[256];
    snprintf(name, sizeof name, "output/output_%06d.%s", step, ext);
    FILE* fp = fopen(name, "w");
    if (!fp) fprintf(stderr, "Error: Could not open file %s for writing.\n", name);
    return fp;
}

// Columns x,y,z (%.6f), rho,u,v,w,p,gx,gy,gz (%.6e); every skip-th cell per
// axis, skip 2 above 128 cells along x and 4 above 256.
void write_csv(const Geometry& g, const HostFields& f, int step) {
    FILE* fp = open_output(step, "csv");
    if (!fp) return;
    fputs("x,y,z,rho,u,v,w,p,gx,gy,gz\n", fp);
    const int skip = g.nx > 256 ? 4 : (g.nx > 128 ? 2 : 1);
    for (int i = 0; i < g.nx; i += skip) {
        const double x = g.xmin + (i + 0.5) * g.dx;
        for (int j = 0; j < g.ny; j += skip) {
            const double y = g.ymin + (j + 0.5) * g.dy;
            for (int k = 0; k < g.nz; k += skip) {
                const double z = g.zmin + (k + 0.5) * g.dz;
                const long c = IDX3D((long)i, (long)j, (long)k, (long)g.ny, (long)g.nz);
                const Primitive3D P = c2p(f.cell(c), g.gamma);
                const double ax = g.gravity ? f.accel(0, c) : 0.0;
                const double ay = g.gravity ? f.accel(1, c) : 0.0;
                const double az = g.gravity ? f.accel(2, c) : 0.0;
                fprintf(fp, "%.6f,%.6f,%.6f,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e\n", x, y, z,
                        P.rho, P.u, P.v, P.w, P.p, ax, ay, az);
            }
        }
    }
    fclose(fp);
}

// Legacy ASCII structured points: density, pressure (%e) and the velocity
// vector, every cell in memory order.
void write_vtk(const Geometry& g, const HostFields& f, int step) {
    FILE* fp = open_output(step, "vtk");
    if (!fp) return;
    const long n = (long)g.nx * g.ny * g.nz;
    fprintf(fp,
            "# vtk DataFile Version 3.0\n3D Hydro Simulation Output\nASCII\n"
            "DATASET STRUCTURED_POINTS\nDIMENSIONS %d %d %d\n",
            g.nx, g.ny, g.nz);
    fprintf(fp, "ORIGIN %f %f %f\n", g.xmin, g.ymin, g.zmin);
    fprintf(fp, "SPACING %f %f %f\n", g.dx, g.dy, g.dz);
    fprintf(fp, "POINT_DATA %d\n", g.nx * g.ny * g.nz);
    fputs("SCALARS density float 1\nLOOKUP_TABLE default\n", fp);
    for (long c = 0; c < n; ++c) fprintf(fp, "%e\n", c2p(f.cell(c), g.gamma).rho);
    fputs("SCALARS pressure float 1\nLOOKUP_TABLE default\n", fp);
    for (long c = 0; c < n; ++c) fprintf(fp, "%e\n", c2p(f.cell(c), g.gamma).p);
    fputs("VECTORS velocity float\n", fp);
    for (long c = 0; c < n; ++c) {
        const Primitive3D P = c2p(f.cell(c), g.gamma);
        fprintf(fp, "%e %e %e\n", P.u, P.v, P.w);
    }
    fclose(fp);
}

HostFields mirror_fields(const GodunovHLLC3D* s) {
    HostFields f;
    f.aos = s->U;
    f.g[0] = s->gx;
    f.g[1] = s->gy;
    f.g[2] = s->gz;
    return f;
}

// ------------------------------------------------------- CSV restart ----
// What a restart file means to the reference (hydro_solver_3d.c:2139-2235):
//  1. The first line is a header; a file that cannot be read, or whose
//     header mentions neither "rho" nor "u" (both must appear), is refused
//     with -1 and U is left alone.
//  2. Otherwise every cell of U is zeroed (g is left as it is).
//  3. Each later line (read in pieces of at most 1023 bytes) holds numbers
//     x, y, z, rho, u, v, w, p and optionally gx, gy, gz, each directly
//     followed by a comma except the last; leading blanks before a number
//     are allowed.  Fewer than 8 leading numbers: the line is skipped.
//  4. The cell is (int) of (x - xmin) / dx truncated toward zero, likewise
//     for y and z; a point outside the grid is skipped.
//  5. U of the cell = prim_to_cons(rho, u, v, w, p), no floors; with gravity
//     on and all 11 numbers present, the cell's g is set as well.
//  6. The accepted points (repeats included) are reported, with a warning
//     when fewer than ncells / 10 (integer division); the result is 0.
struct RestartRow {
    double v[11];
    int count;  // leading numbers read (<= 11)
};

RestartRow parse_row(const char* text) {
    RestartRow row{};
    const char* p = text;
    while (row.count < 11) {
        char* end = nullptr;
        const double x = strtod(p, &end);
        if (end == p) break;  // no number here
        row.v[row.count++] = x;
        if (*end != ',') break;
        p = end + 1;
    }
    return row;
}

// Truncated cell coordinate of q = (x - min) / h in [0, n): false outside.
bool cell_coord(double q, int n, int* out) {
    if (!(q > -1.0) || !(q < (double)n)) return false;
    *out = (int)q;  // toward zero, so (-1, 0) maps to cell 0
    return true;
}

int read_csv_restart(GodunovHLLC3D* s, FILE* fp, const char* path) {
    char line[1024];
    if (!fgets(line, sizeof line, fp)) {
        fprintf(stderr, "Error: restart file '%s' is empty\n", path);
        return -1;
    }
    if (!strstr(line, "rho") || !strstr(line, "u")) {
        fprintf(stderr,
                "Error: restart file '%s' has no x,y,z,rho,u,v,w,p[,gx,gy,gz] header line\n",
                path);
        return -1;
    }
    const long ncells = (long)s->nx * s->ny * s->nz;
    memset(s->U, 0, sizeof(State3D) * (size_t)ncells);
    long accepted = 0;
    while (fgets(line, sizeof line, fp)) {
        const RestartRow row = parse_row(line);
        if (row.count < 8) continue;
        int i, j, k;
        if (!cell_coord((row.v[0] - s->xmin) / s->dx, s->nx, &i) ||
            !cell_coord((row.v[1] - s->ymin) / s->dy, s->ny, &j) ||
            !cell_coord((row.v[2] - s->zmin) / s->dz, s->nz, &k))
            continue;
        const long c = IDX3D((long)i, (long)j, (long)k, (long)s->ny, (long)s->nz);
        const Primitive3D P = {row.v[3], row.v[4], row.v[5], row.v[6], row.v[7]};
        s->U[c] = p2c(P, s->gamma);
        if (s->enable_gravity && row.count >= 11) {
            s->gx[c] = row.v[8];
            s->gy[c] = row.v[9];
            s->gz[c] = row.v[10];
        }
        ++accepted;
    }
    printf("Restart file %s: %ld of %ld cells set\n", path, accepted, ncells);
    if (accepted < ncells / 10)
        fprintf(stderr,
                "Warning: the restart file set fewer than a tenth of the cells (%ld of %ld); "
                "it may be subsampled, from another grid, or cut short\n",
                accepted, ncells);
    return 0;
}

// ------------------------------------------------------ binary files ----
// HYDIC001: magic, int32 nx ny nz, then U (AoS doubles).
// HYDCK001: magic, int32 nx ny nz, int32 flags (1: potential, 2: g),
//           double t, int32 output count, int32 0, then U, the potential
//           (natural order) and gx, gy, gz planes as the flags say.
constexpr char kIcMagic[8] = {'H', 'Y', 'D', 'I', 'C', '0', '0', '1'};
constexpr char kCkMagic[8] = {'H', 'Y', 'D', 'C', 'K', '0', '0', '1'};
enum { CK_PHI = 1, CK_G = 2 };

struct CkHeader {
    char magic[8];
    int32_t nx, ny, nz, flags;
    double t;
    int32_t output_count, reserved;
};
static_assert(sizeof(CkHeader) == 40, "checkpoint header is 40 bytes");

bool read_exact(FILE* fp, void* dst, size_t bytes) { return fread(dst, 1, bytes, fp) == bytes; }

// Reads a binary state into s (U, and the potential and g when the file has
// them and gravity is on); s is changed only when the whole file is good.
int read_binary_state(GodunovHLLC3D* s, FILE* fp, const char* path, double* t, int* count) {
    char magic[8];
    if (!read_exact(fp, magic, 8)) {
        fprintf(stderr, "Error: '%s' is empty\n", path);
        return -1;
    }
    const bool ck = memcmp(magic, kCkMagic, 8) == 0;
    if (!ck && memcmp(magic, kIcMagic, 8) != 0) {
        fprintf(stderr, "Error: '%s' is neither an initial-state nor a checkpoint file\n", path);
        return -1;
    }
    CkHeader h{};
    memcpy(h.magic, magic, 8);
    const size_t rest = ck ? sizeof(CkHeader) - 8 : 3 * sizeof(int32_t);
    if (!read_exact(fp, &h.nx, rest) || h.nx != s->nx || h.ny != s->ny || h.nz != s->nz) {
        fprintf(stderr, "Error: '%s' does not hold a %d x %d x %d grid\n", path, s->nx, s->ny,
                s->nz);
        return -1;
    }
    const long n = (long)s->nx * s->ny * s->nz;
    std::vector<State3D> U((size_t)n);
    std::vector<double> phi, g;
    bool ok = read_exact(fp, U.data(), sizeof(State3D) * (size_t)n);
    if (ok && ck && (h.flags & CK_PHI)) {
        phi.resize((size_t)n);
        ok = read_exact(fp, phi.data(), sizeof(double) * (size_t)n);
    }
    if (ok && ck && (h.flags & CK_G)) {
        g.resize(3 * (size_t)n);
        ok = read_exact(fp, g.data(), 3 * sizeof(double) * (size_t)n);
    }
    if (!ok) {
        fprintf(stderr, "Error: '%s' is truncated\n", path);
        return -1;
    }
    memcpy(s->U, U.data(), sizeof(State3D) * (size_t)n);
    s->dev->uploaded = false;
    GravitySolver3D* gs = s->enable_gravity ? s->gravity_solver : nullptr;
    if (gs && (!phi.empty() || !g.empty())) {
        if (!phi.empty()) memcpy(gs->phi, phi.data(), sizeof(double) * (size_t)n);
        if (!g.empty())
            for (int c = 0; c < 3; ++c) {
                double* dst = c == 0 ? s->gx : (c == 1 ? s->gy : s->gz);
                memcpy(dst, g.data() + c * n, sizeof(double) * (size_t)n);
            }
        if (grav_internal_restore(gs, phi.empty() ? nullptr : phi.data(),
                                  g.empty() ? nullptr : g.data(),
                                  g.empty() ? nullptr : g.data() + n,
                                  g.empty() ? nullptr : g.data() + 2 * n))
            return -2;
    }
    if (t) *t = ck ? h.t : 0.0;
    if (count) *count = ck ? h.output_count : 0;
    return 0;
}

bool has_binary_magic(FILE* fp) {
    char magic[8];
    const bool got = read_exact(fp, magic, 8);
    rewind(fp);
    return got && (memcmp(magic, kCkMagic, 8) == 0 || memcmp(magic, kIcMagic, 8) == 0);
}

}  // namespace

// --------------------------------------------------------- output pipe ----
// One snapshot in flight: device copies of the five state planes and g, the
// pinned host planes they are copied into on `copy` (a second stream, default
// priority), and the thread that writes the files once the copy is done.
struct OutputPipe {
    long n = 0;
    int planes = 0;  // 5, or 8 with g
    hipStream_t copy = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    double* dev[8] = {nullptr};
    double* host[8] = {nullptr};
    std::thread writer;
};

void hyd_output_drain(HydroDevice* d) {
    if (d->out && d->out->writer.joinable()) d->out->writer.join();
}

void hyd_output_free(HydroDevice* d) {
    OutputPipe* p = d->out;
    if (!p) return;
    hyd_output_drain(d);
    for (int f = 0; f < 8; ++f) {
        if (p->dev[f]) (void)hipFree(p->dev[f]);
        if (p->host[f]) (void)hipHostFree(p->host[f]);
    }
    if (p->ready) (void)hipEventDestroy(p->ready);
    if (p->done) (void)hipEventDestroy(p->done);
    if (p->copy) (void)hipStreamDestroy(p->copy);
    delete p;
    d->out = nullptr;
}

static int pipe_create(GodunovHLLC3D* s) {
    HydroDevice* d = s->dev;
    OutputPipe* p = new OutputPipe();
    d->out = p;
    p->n = d->n;
    p->planes = (s->enable_gravity && s->gravity_solver) ? 8 : 5;
    HYD_CHECK(hipStreamCreateWithFlags(&p->copy, hipStreamNonBlocking));
    HYD_CHECK(hipEventCreateWithFlags(&p->ready, hipEventDisableTiming));
    HYD_CHECK(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
    for (int f = 0; f < p->planes; ++f) {
        HYD_CHECK(hipMalloc((void**)&p->dev[f], sizeof(double) * p->n));
        HYD_CHECK(hipHostMalloc((void**)&p->host[f], sizeof(double) * p->n, hipHostMallocDefault));
    }
    return 0;
}

// Queue the CSV and VTK files of output `count` from the current device
// state.  The steps that follow are not held up: the snapshot is a device
// copy on the step stream, its transfer runs on the copy stream, and the
// files are written by a host thread (the previous one is joined first).
int hyd_output_async(GodunovHLLC3D* s, int count) {
    HydroDevice* d = s->dev;
    if (!d->out && pipe_create(s)) {
        hyd_output_free(d);
        return -2;
    }
    OutputPipe* p = d->out;
    hyd_output_drain(d);
    const size_t b = sizeof(double) * p->n;
    for (int f = 0; f < 5; ++f)
        HYD_CHECK(hipMemcpyAsync(p->dev[f], d->U[d->cur][f], b, hipMemcpyDeviceToDevice, d->stream));
    for (int c = 0; c < p->planes - 5; ++c)
        HYD_CHECK(hipMemcpyAsync(p->dev[5 + c], grav_internal_g(s->gravity_solver, c), b,
                                 hipMemcpyDeviceToDevice, d->stream));
    HYD_CHECK(hipEventRecord(p->ready, d->stream));
    HYD_CHECK(hipStreamWaitEvent(p->copy, p->ready, 0));
    for (int f = 0; f < p->planes; ++f)
        HYD_CHECK(hipMemcpyAsync(p->host[f], p->dev[f], b, hipMemcpyDeviceToHost, p->copy));
    HYD_CHECK(hipEventRecord(p->done, p->copy));
    HostFields fields;
    for (int f = 0; f < 5; ++f) fields.plane[f] = p->host[f];
    for (int c = 0; c < p->planes - 5; ++c) fields.g[c] = p->host[5 + c];
    const Geometry geo = geometry(s);
    hipEvent_t done = p->done;
    p->writer = std::thread([geo, fields, count, done]() {
        if (hipEventSynchronize(done) != hipSuccess) {
            fprintf(stderr, "hydro_mi355x: output %d lost (device copy failed)\n", count);
            return;
        }
        write_csv(geo, fields, count);
        write_vtk(geo, fields, count);
    });
    return 0;
}

int hyd_read_state_file(GodunovHLLC3D* s, const char* path, double* t, int* count) {
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        fprintf(stderr, "Error: cannot open '%s': %s\n", path, strerror(errno));
        return -1;
    }
    const int rc = read_binary_state(s, fp, path, t, count);
    fclose(fp);
    return rc;
}

extern "C" {

void write_output_to_file_3d(const GodunovHLLC3D* s, int step) {
    write_csv(geometry(s), mirror_fields(s), step);
}

void write_output_vtk(const GodunovHLLC3D* s, int step) {
    write_vtk(geometry(s), mirror_fields(s), step);
}

// The reference's CSV restart (semantics above); a checkpoint or initial-
// state file of this build is recognised by its first bytes and loaded at
// full precision instead, potential and g included.
int load_from_csv_restart(GodunovHLLC3D* s, const char* filename) {
    FILE* fp = fopen(filename, "rb");
    if (!fp) {
        fprintf(stderr, "Error: cannot open restart file '%s': %s\n", filename, strerror(errno));
        return -1;
    }
    int rc;
    if (has_binary_magic(fp)) {
        rc = read_binary_state(s, fp, filename, nullptr, nullptr);
        if (rc == 0) printf("Restart file %s: checkpoint loaded\n", filename);
    } else {
        rc = read_csv_restart(s, fp, filename);
        s->dev->uploaded = false;
    }
    fclose(fp);
    if (rc == 0) s->ic_status = 0;
    return rc < 0 ? -1 : rc;
}

int solver3d_write_checkpoint(GodunovHLLC3D* s, const char* path, double t, int output_count) {
    if (hyd_download(s)) return -2;
    GravitySolver3D* gs = (s->enable_gravity && s->gravity_solver) ? s->gravity_solver : nullptr;
    if (gs && gravity_sync_to_host(gs)) return -2;
    FILE* fp = fopen(path, "wb");
    if (!fp) {
        fprintf(stderr, "Error: cannot write checkpoint '%s': %s\n", path, strerror(errno));
        return -1;
    }
    CkHeader h{};
    memcpy(h.magic, kCkMagic, 8);
    h.nx = s->nx;
    h.ny = s->ny;
    h.nz = s->nz;
    h.flags = gs ? (CK_PHI | CK_G) : 0;
    h.t = t;
    h.output_count = output_count;
    const size_t n = (size_t)s->nx * s->ny * s->nz;
    bool ok = fwrite(&h, sizeof h, 1, fp) == 1 && fwrite(s->U, sizeof(State3D), n, fp) == n;
    if (ok && gs)
        ok = fwrite(gs->phi, sizeof(double), n, fp) == n &&
             fwrite(gs->gx, sizeof(double), n, fp) == n &&
             fwrite(gs->gy, sizeof(double), n, fp) == n &&
             fwrite(gs->gz, sizeof(double), n, fp) == n;
    if (fclose(fp) != 0) ok = false;
    if (!ok) {
        fprintf(stderr, "Error: checkpoint '%s' could not be written in full\n", path);
        return -1;
    }
    return 0;
}

int solver3d_read_checkpoint(GodunovHLLC3D* s, const char* path, double* t, int* output_count) {
    const int rc = hyd_read_state_file(s, path, t, output_count);
    if (rc == 0) s->ic_status = 0;
    return rc;
}

}  // extern "C"
