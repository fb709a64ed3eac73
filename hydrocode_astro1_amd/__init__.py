"""MI355X-native hot path of the hydrocode_astro1 solver (HLLC Godunov hydro,
CFL dt, red-black SOR Poisson with a multipole boundary) behind the
reference's C API.

The product is the shared library libhydro_mi355x.so (HIP kernels for gfx950
and the C ABI declared in include/).  This module loads it with ctypes and
mirrors the reference interface for Python callers and tests:

  * Solver   -- GodunovHLLC3D: solver3d_create / set state / step / run
  * Gravity  -- GravitySolver3D: gravity_solver_create / gravity_solve / ...
  * InputDeck helpers -- input_deck_create and solver3d_create_from_deck

Loading fails loudly when the library has not been built; there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

from . import ics  # noqa: F401  (initial conditions for the benchmark configs)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HYDRO_MI355X_LIB") or os.path.join(_HERE, "libhydro_mi355x.so")

BC_PERIODIC, BC_REFLECTIVE, BC_OUTFLOW = 0, 1, 2
GRAV_BC_PERIODIC, GRAV_BC_ISOLATED, GRAV_BC_DIRICHLET = 0, 1, 2
POISSON_FP64_SOR, POISSON_FP32_DC, POISSON_NN_FP32_DC = 0, 1, 2
SOR_CRIT_MAX_UPDATE, SOR_CRIT_REL_RESID = 0, 1
DC_INNER_SOR, DC_INNER_MULTIGRID = 0, 1
MOMENT_ORDER_TREE, MOMENT_ORDER_SEQUENTIAL, MOMENT_ORDER_AUTO = 0, 1, 2
# ProblemType (input_deck.h): the values the reference has generators for
(KELVIN_HELMHOLTZ_3D, BLAST_WAVE_3D, STELLAR_PROFILE_3D, BINARY_STAR_3D,
 STELLAR_OCTANT_EXPLOSION_3D) = 11, 12, 15, 16, 17

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class GravitySolver3D(C.Structure):
    """Layout of GravitySolver3D (include/gravity_solver_3d.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("G", C.c_double),
        ("phi", _dp), ("rho", _dp), ("gx", _dp), ("gy", _dp), ("gz", _dp),
        ("omega", C.c_double), ("max_iterations", C.c_int), ("tolerance", C.c_double),
        ("bc_type", C.c_int),
        ("last_iterations", C.c_int), ("last_residual", C.c_double),
        ("dev", C.c_void_p),
        ("multipole_lmax", C.c_int), ("poisson_mode", C.c_int), ("criterion", C.c_int),
        ("last_status", C.c_int),
        ("last_rel_residual", C.c_double), ("last_solve_ms", C.c_double),
        ("nn_model", C.c_void_p), ("last_guess", C.c_int),
        ("last_guess_rel_residual", C.c_double), ("last_warm_rel_residual", C.c_double),
        ("dc_inner", C.c_int), ("moment_order", C.c_int),
        ("nn_correction", C.c_void_p), ("last_corr_rel_residual", C.c_double),
    ]


class GodunovHLLC3D(C.Structure):
    """Layout of GodunovHLLC3D (include/hydro_solver_3d.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
        ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dt", C.c_double), ("cfl", C.c_double),
        ("bc_xmin", C.c_int), ("bc_xmax", C.c_int), ("bc_ymin", C.c_int),
        ("bc_ymax", C.c_int), ("bc_zmin", C.c_int), ("bc_zmax", C.c_int),
        ("problem", C.c_int), ("gamma", C.c_double),
        ("U", _dp), ("U_half", _dp), ("F_x", _dp), ("F_y", _dp), ("F_z", _dp),
        ("gx", _dp), ("gy", _dp), ("gz", _dp),
        ("gravity_solver", C.POINTER(GravitySolver3D)),
        ("enable_gravity", C.c_bool), ("G_const", C.c_double),
        ("dev", C.c_void_p), ("max_steps", C.c_int), ("output_enable", C.c_bool),
        ("steps_taken", C.c_long), ("ic_status", C.c_int), ("checkpoint_enable", C.c_bool),
    ]


_SP = C.POINTER(GodunovHLLC3D)
_GP = C.POINTER(GravitySolver3D)

# (name, restype, argtypes) of every exported symbol the headers declare
EXPORTS = [
    ("solver3d_create", _SP, [C.c_int] * 3 + [C.c_double] * 7),
    ("solver3d_destroy", None, [_SP]),
    ("set_initial_conditions_3d", None, [_SP, C.c_int]),
    ("set_initial_conditions_3d_from_deck", None, [_SP, C.c_void_p]),
    ("run_simulation_3d", None, [_SP, C.c_double, C.c_double]),
    ("run_simulation_3d_restart", None, [_SP, C.c_double, C.c_double, C.c_double, C.c_int]),
    ("write_output_to_file_3d", None, [_SP, C.c_int]),
    ("write_output_vtk", None, [_SP, C.c_int]),
    ("load_from_csv_restart", C.c_int, [_SP, C.c_char_p]),
    ("solver3d_create_from_deck", _SP, [C.c_void_p]),
    ("gravity_solver_create", _GP, [C.c_int] * 3 + [C.c_double] * 7),
    ("gravity_solver_destroy", None, [_GP]),
    ("gravity_solve_potential", C.c_int, [_GP, _dp]),
    ("gravity_compute_acceleration", None, [_GP]),
    ("gravity_set_solver_params", None, [_GP, C.c_double, C.c_int, C.c_double]),
    ("gravity_set_boundary_type", None, [_GP, C.c_int]),
    ("gravity_get_diagnostics", None, [_GP, _ip, _dp]),
    ("gravity_solve", C.c_int, [_GP, _dp]),
    ("gravity_write_potential", None, [_GP, C.c_char_p]),
    ("input_deck_create", C.c_void_p, [C.c_char_p]),
    ("input_deck_destroy", None, [C.c_void_p]),
    ("input_deck_print", None, [C.c_void_p]),
    ("parse_problem_type", C.c_int, [C.c_char_p]),
    ("parse_bc_type", C.c_int, [C.c_char_p]),
    ("hydro_device_count", C.c_int, []),
    ("solver3d_upload", C.c_int, [_SP]),
    ("solver3d_download", C.c_int, [_SP]),
    ("solver3d_set_state", C.c_int, [_SP, _dp]),
    ("solver3d_write_checkpoint", C.c_int, [_SP, C.c_char_p, C.c_double, C.c_int]),
    ("solver3d_read_checkpoint", C.c_int, [_SP, C.c_char_p, _dp, _ip]),
    ("solver3d_step", C.c_int, [_SP]),
    ("solver3d_run_steps", C.c_int, [_SP, C.c_int, _dp]),
    ("solver3d_run", C.c_int, [_SP, C.c_double, C.c_int, _dp, _dp, _ip, C.c_int]),
    ("solver3d_last_timing", C.c_int, [_SP, _dp, _dp]),
    ("solver3d_benchmark_steps", C.c_int, [_SP, C.c_int, _dp, _dp, _dp]),
    ("hydro_sequential_sum", C.c_double, [_dp, C.c_long, C.c_int, C.POINTER(C.c_longlong), _ip]),
    ("hydro_fill_initial_state", C.c_int, [C.c_int, C.c_int, C.c_int] + [C.c_double] * 8
     + [C.c_int, C.c_void_p, _dp]),
    ("gravity_set_multipole_lmax", C.c_int, [_GP, C.c_int]),
    ("gravity_set_poisson_mode", C.c_int, [_GP, C.c_int]),
    ("gravity_set_dc_inner", C.c_int, [_GP, C.c_int]),
    ("gravity_set_moment_order", C.c_int, [_GP, C.c_int]),
    ("gravity_set_criterion", C.c_int, [_GP, C.c_int]),
    ("gravity_sync_to_device", C.c_int, [_GP]),
    ("gravity_sync_to_host", C.c_int, [_GP]),
    ("gravity_moments", C.c_int, [_GP, _dp, _dp, _dp]),
    ("gravity_boundary_potential", C.c_int, [_GP, _dp, _dp]),
    ("gravity_rel_residual", C.c_int, [_GP, _dp, _dp]),
    ("gravity_solve_device", C.c_int, [_GP, C.c_void_p]),
    ("gravity_moments_noalloc_scratch", C.c_long, [C.c_int] * 4),
    ("gravity_moments_noalloc", C.c_int,
     [C.c_int] * 3 + [C.c_double] * 6 + [C.c_int, C.c_void_p, C.c_void_p]),
    ("gravity_set_nn_model", C.c_int, [_GP, C.c_void_p]),
    ("gravity_set_nn_correction", C.c_int, [_GP, C.c_void_p]),
    ("nn_model_create", C.c_void_p, [C.c_ulonglong]),
    ("nn_model_destroy", None, [C.c_void_p]),
    ("nn_model_num_params", C.c_long, [C.c_void_p]),
    ("nn_model_param_len", C.c_long, [C.c_void_p]),
    ("nn_model_param_offset", C.c_long, [C.c_void_p, C.c_int, C.c_int]),
    ("nn_model_save", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nn_model_load", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nn_model_get_params", C.c_int, [C.c_void_p, _fp]),
    ("nn_model_set_params", C.c_int, [C.c_void_p, _fp]),
    ("nn_model_set_kind", C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    ("nn_model_kind", C.c_int, [C.c_void_p, _dp]),
    ("nn_model_forward_backward", C.c_int,
     [C.c_void_p, _fp, _fp, C.c_int, _dp, C.c_double, _fp, _fp, _fp, _fp]),
    ("nn_dataset_create", C.c_void_p, [C.c_int, C.c_int]),
    ("nn_dataset_destroy", None, [C.c_void_p]),
    ("nn_dataset_fill", C.c_int,
     [C.c_void_p, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_double, _GP]),
    ("nn_dataset_get", C.c_int, [C.c_void_p, C.c_int, _fp, _fp, _dp]),
    ("nn_dataset_set", C.c_int, [C.c_void_p, C.c_int, _fp, _fp]),
    ("nn_train", C.c_int, [C.c_void_p, C.c_void_p, _ip, C.c_int, C.c_int, C.c_float, _fp]),
    ("nn_evaluate", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _fp]),
    ("nn_benchmark", C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp]),
    ("nn_forward_flops", C.c_double, [C.c_int]),
]

_lib = None


def lib():
    """The loaded device library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with python -m hydrocode_astro1_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def initial_state(problem, shape, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), gamma=5.0 / 3.0, G=1.0,
                  deck=None):
    """The reference's initial-condition problem as an (nx, ny, nz, 5) array
    of conserved variables, made on the host by hydro_fill_initial_state:
    `problem` alone as set_initial_conditions_3d, or `deck` (a deck file
    path, whose problem and parameters are used) as
    set_initial_conditions_3d_from_deck.  None when the problem has no
    generator."""
    nx, ny, nz = shape
    U = np.zeros((nx, ny, nz, 5))
    L = lib()
    dp = L.input_deck_create(deck.encode()) if deck else None
    try:
        rc = L.hydro_fill_initial_state(nx, ny, nz, *box, gamma, G, int(problem), dp, _d(U))
    finally:
        if dp:
            L.input_deck_destroy(dp)
    if rc < 0:
        raise ValueError("hydro_fill_initial_state: bad arguments")
    return U if rc == 0 else None


SEQSUM_STATS = ("chunks", "chunks_p1_3", "chunks_p4_6", "chunks_added", "verify_failed",
                "cycles_stitch", "groups_walked", "chain_redo", "chunks_zero")


def sequential_sum(x, method=1, stats=False):
    """((0 + x[0]) + x[1]) + ... rounded after every addition, computed on
    the device: method 0 adds the terms in one chain, method 1 by chunks
    with tabulated, certified increments carried group by group (the moment
    sums of MOMENT_ORDER_SEQUENTIAL use it), method 2 as 1 and then re-runs
    every chunk from its exact start (a check of the certificates), method 3
    as 1 with the tables built by the LDS-staged kernel that the moment sums
    use (a single sum otherwise takes the per-lane kernel).
    stats=True also returns methods 1, 2 and 3's counters as a dict."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    st = C.c_int(-1)
    cnt = (C.c_longlong * 9)()
    r = lib().hydro_sequential_sum(_d(x), x.size, method, cnt if stats else None, C.byref(st))
    _check(st.value, "hydro_sequential_sum")
    return (r, dict(zip(SEQSUM_STATS, list(cnt)))) if stats else r


def _check(rc, what):
    if rc is not None and rc < -1:
        raise RuntimeError(f"{what} failed with device error {rc}")
    return rc


class Gravity:
    """GravitySolver3D on the device (gravity_solver_3d.h)."""

    def __init__(self, nx, ny, nz, box=(0, 1, 0, 1, 0, 1), G=1.0, handle=None):
        self.L = lib()
        self.owned = handle is None
        if handle is None:
            handle = self.L.gravity_solver_create(nx, ny, nz, *[float(b) for b in box], float(G))
            if not handle:
                raise RuntimeError("gravity_solver_create failed")
        self.h = handle
        self.shape = (handle.contents.nx, handle.contents.ny, handle.contents.nz)
        self.n = int(np.prod(self.shape))

    def close(self):
        if self.owned and getattr(self, "h", None):
            self.L.gravity_solver_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def s(self):
        return self.h.contents

    def set_params(self, omega=1.5, max_iter=10000, tol=1e-6):
        self.L.gravity_set_solver_params(self.h, omega, max_iter, tol)

    def set_boundary(self, bc):
        self.L.gravity_set_boundary_type(self.h, bc)

    def set_lmax(self, lmax):
        if self.L.gravity_set_multipole_lmax(self.h, lmax) != 0:
            raise ValueError(f"bad lmax {lmax}")

    def set_mode(self, mode):
        if self.L.gravity_set_poisson_mode(self.h, mode) != 0:
            raise ValueError(f"bad mode {mode}")

    def set_dc_inner(self, inner):
        """DC_INNER_SOR or DC_INNER_MULTIGRID for the FP32 modes."""
        if self.L.gravity_set_dc_inner(self.h, inner) != 0:
            raise ValueError(f"bad inner solver {inner} for this grid")

    def set_moment_order(self, order):
        """MOMENT_ORDER_AUTO (default: the reference's cell-after-cell sums
        under FP64 SOR, the tree order in the FP32 modes), MOMENT_ORDER_TREE
        or MOMENT_ORDER_SEQUENTIAL."""
        if self.L.gravity_set_moment_order(self.h, order) != 0:
            raise ValueError(f"bad moment order {order}")

    def set_nn_model(self, model):
        """Attach an nn.Model for POISSON_NN_FP32_DC (None detaches)."""
        if self.L.gravity_set_nn_model(self.h, model.h if model is not None else None) != 0:
            raise ValueError("the network is not a potential guess")
        self._nn = model

    def set_nn_correction(self, model):
        """Attach a start-correction nn.Model (kind 1) for POISSON_NN_FP32_DC
        (None detaches)."""
        if self.L.gravity_set_nn_correction(self.h,
                                            model.h if model is not None else None) != 0:
            raise ValueError("the network is not a start correction")
        self._nn_corr = model

    def set_criterion(self, crit):
        if self.L.gravity_set_criterion(self.h, crit) != 0:
            raise ValueError(f"bad criterion {crit}")

    def _rho(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert r.size == self.n
        return r

    def solve(self, rho):
        """gravity_solve: potential, then (when converged) acceleration."""
        self._keep = self._rho(rho)
        return _check(self.L.gravity_solve(self.h, _d(self._keep)), "gravity_solve")

    def solve_potential(self, rho):
        self._keep = self._rho(rho)
        return _check(self.L.gravity_solve_potential(self.h, _d(self._keep)),
                      "gravity_solve_potential")

    def compute_acceleration(self):
        self.L.gravity_compute_acceleration(self.h)

    def diagnostics(self):
        it = C.c_int(0)
        res = C.c_double(0.0)
        self.L.gravity_get_diagnostics(self.h, C.byref(it), C.byref(res))
        return it.value, res.value

    def phi(self):
        return np.ctypeslib.as_array(self.s.phi, shape=(self.n,)).reshape(self.shape).copy()

    def set_phi(self, phi):
        a = np.ctypeslib.as_array(self.s.phi, shape=(self.n,))
        a[:] = np.asarray(phi, dtype=np.float64).reshape(-1)
        _check(self.L.gravity_sync_to_device(self.h), "gravity_sync_to_device")

    def g(self):
        return [np.ctypeslib.as_array(getattr(self.s, c), shape=(self.n,)).reshape(self.shape).copy()
                for c in ("gx", "gy", "gz")]

    def moments(self, rho):
        out = np.zeros(13)
        hi = np.zeros(256)
        _check(self.L.gravity_moments(self.h, _d(self._rho(rho)), _d(out), _d(hi)),
               "gravity_moments")
        return out, hi

    def boundary_potential(self, rho):
        out = np.zeros(self.n)
        _check(self.L.gravity_boundary_potential(self.h, _d(self._rho(rho)), _d(out)),
               "gravity_boundary_potential")
        return out.reshape(self.shape)

    def rel_residual(self, rho):
        r = C.c_double(0.0)
        if _check(self.L.gravity_rel_residual(self.h, _d(self._rho(rho)), C.byref(r)),
                  "gravity_rel_residual") == -1:
            raise ValueError("gravity_rel_residual is not provided for nz == 1")
        return r.value


class Solver:
    """GodunovHLLC3D on the device (hydro_solver_3d.h)."""

    def __init__(self, nx, ny, nz, box=(0, 1, 0, 1, 0, 1), G=0.0, gamma=5.0 / 3.0, cfl=0.4,
                 bc=(BC_OUTFLOW,) * 6, handle=None):
        self.L = lib()
        if handle is None:
            handle = self.L.solver3d_create(nx, ny, nz, *[float(b) for b in box], float(G))
            if not handle:
                raise RuntimeError("solver3d_create failed")
            s = handle.contents
            s.gamma = gamma
            s.cfl = cfl
            (s.bc_xmin, s.bc_xmax, s.bc_ymin, s.bc_ymax, s.bc_zmin, s.bc_zmax) = bc
        self.h = handle
        s = self.h.contents
        self.shape = (s.nx, s.ny, s.nz)
        self.n = int(np.prod(self.shape))
        self.gravity = (Gravity(0, 0, 0, handle=s.gravity_solver)
                        if s.enable_gravity and s.gravity_solver else None)

    def close(self):
        if getattr(self, "h", None):
            self.gravity = None
            self.L.solver3d_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def s(self):
        return self.h.contents

    @property
    def dt(self):
        return self.s.dt

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        assert U.size == 5 * self.n
        _check(self.L.solver3d_set_state(self.h, _d(U)), "solver3d_set_state")

    def state(self):
        _check(self.L.solver3d_download(self.h), "solver3d_download")
        return np.ctypeslib.as_array(self.s.U, shape=(5 * self.n,)).reshape(
            self.shape + (5,)).copy()

    def write_checkpoint(self, path, t=0.0, output_count=0):
        rc = self.L.solver3d_write_checkpoint(self.h, str(path).encode(), float(t),
                                              int(output_count))
        if rc != 0:
            raise RuntimeError(f"solver3d_write_checkpoint({path}) failed with {rc}")

    def read_checkpoint(self, path):
        """Load a checkpoint; returns (t, output_count) from its header."""
        t, n = C.c_double(0.0), C.c_int(0)
        rc = self.L.solver3d_read_checkpoint(self.h, str(path).encode(), C.byref(t), C.byref(n))
        if rc != 0:
            raise RuntimeError(f"solver3d_read_checkpoint({path}) failed with {rc}")
        return t.value, n.value

    def g(self):
        _check(self.L.solver3d_download(self.h), "solver3d_download")
        return [np.ctypeslib.as_array(getattr(self.s, c), shape=(self.n,)).reshape(self.shape).copy()
                for c in ("gx", "gy", "gz")]

    def step(self):
        return _check(self.L.solver3d_step(self.h), "solver3d_step")

    def run_steps(self, n):
        log = np.zeros(max(n, 1))
        _check(self.L.solver3d_run_steps(self.h, n, _d(log)), "solver3d_run_steps")
        return log[:n]

    def run(self, t_end, max_steps=0, cap=100000):
        t = C.c_double(0.0)
        dts = np.zeros(cap)
        its = np.zeros(cap, dtype=np.int32)
        n = self.L.solver3d_run(self.h, t_end, max_steps, C.byref(t), _d(dts),
                                its.ctypes.data_as(_ip), cap)
        if n < 0:
            raise RuntimeError(f"solver3d_run failed with {n}")
        return n, t.value, dts[:n].copy(), its[:n].copy()

    def benchmark(self, nsteps):
        """(wall ms of nsteps steps, avg step-kernel ms, avg boundary-kernel ms)"""
        w, a, b = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
        _check(self.L.solver3d_benchmark_steps(self.h, nsteps, C.byref(w), C.byref(a),
                                               C.byref(b)), "solver3d_benchmark_steps")
        return w.value, a.value, b.value

    def timing(self):
        a, b = C.c_double(0.0), C.c_double(0.0)
        _check(self.L.solver3d_last_timing(self.h, C.byref(a), C.byref(b)), "timing")
        return a.value, b.value


def device_count():
    return lib().hydro_device_count()
