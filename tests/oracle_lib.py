"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, when it was
built from the reference sources, to the compiled reference
(oracle/_ref/libref.so).  Test infrastructure: only tests, smoke() and the
cpu_baseline leg of bench.py use this module."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


BC_PERIODIC, BC_REFLECTIVE, BC_OUTFLOW = 0, 1, 2
MOM_SEQUENTIAL, MOM_DDTREE = 0, 1
GRAV_PERIODIC, GRAV_ISOLATED, GRAV_DIRICHLET = 0, 1, 2


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_olib = None
_rlib = None


def oracle_lib():
    global _olib
    if _olib is None:
        lib = _load(ORACLE_SO)
        if lib is None:
            raise RuntimeError("oracle/liboracle.so missing: run make -C oracle")
        lib.or_create.restype = C.c_void_p
        lib.or_create.argtypes = [C.c_int] * 3 + [C.c_double] * 7
        lib.or_destroy.argtypes = [C.c_void_p]
        lib.or_set_hydro.argtypes = [C.c_void_p, C.c_double, C.c_double, _ip]
        lib.or_set_gravity.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double,
                                       C.c_int, C.c_int, C.c_int, C.c_int]
        for f in ("or_set_state", "or_get_state", "or_set_phi", "or_get_phi"):
            getattr(lib, f).argtypes = [C.c_void_p, _dp]
        lib.or_get_g.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.or_get_grav_g.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.or_get_dt.restype = C.c_double
        lib.or_get_dt.argtypes = [C.c_void_p]
        lib.or_get_diag.argtypes = [C.c_void_p, _ip, _dp, _ip]
        lib.or_step.argtypes = [C.c_void_p]
        lib.or_run.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _ip]
        lib.or_gravity_solve.argtypes = [C.c_void_p, _dp]
        lib.or_moments.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.or_boundary_phi.argtypes = [C.c_void_p, _dp, _dp]
        lib.or_rel_residual.restype = C.c_double
        lib.or_rel_residual.argtypes = [C.c_void_p, _dp, _dp]
        lib.or_solid_harmonics.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        lib.or_ddtree_sum.restype = C.c_double
        lib.or_ddtree_sum.argtypes = [_dp, C.c_long]
        lib.or_cons_to_prim_batch.argtypes = [C.c_int, _dp, C.c_double, _dp]
        lib.or_prim_to_cons_batch.argtypes = [C.c_int, _dp, C.c_double, _dp]
        lib.or_hllc_batch.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]
        lib.or_compute_dt.restype = C.c_double
        lib.or_compute_dt.argtypes = [_dp, C.c_int, C.c_int, C.c_int] + [C.c_double] * 5
        _olib = lib
    return _olib


def ref_lib():
    """The compiled reference, or None when it was not built."""
    global _rlib
    if _rlib is None:
        lib = _load(REF_SO)
        if lib is None:
            return None
        lib.ref_create.restype = C.c_void_p
        lib.ref_create.argtypes = [C.c_int] * 3 + [C.c_double] * 7
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_set_hydro.argtypes = [C.c_void_p, C.c_double, C.c_double, _ip]
        lib.ref_set_grav.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
        lib.ref_set_state.argtypes = [C.c_void_p, _dp]
        lib.ref_get_state.argtypes = [C.c_void_p, _dp]
        lib.ref_get_g.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.ref_get_phi.argtypes = [C.c_void_p, _dp]
        lib.ref_diag.argtypes = [C.c_void_p, _ip, _dp]
        lib.ref_dt.restype = C.c_double
        lib.ref_dt.argtypes = [C.c_void_p]
        lib.ref_step.argtypes = [C.c_void_p]
        lib.ref_run.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _ip]
        lib.ref_ic_problem.argtypes = [C.c_void_p, C.c_int]
        lib.ref_ic_deck.argtypes = [C.c_void_p, C.c_char_p]
        lib.ref_set_gamma.argtypes = [C.c_void_p, C.c_double]
        lib.ref_from_deck.restype = C.c_void_p
        lib.ref_from_deck.argtypes = [C.c_char_p]
        lib.ref_c2p_batch.argtypes = [C.c_int, _dp, C.c_double, _dp]
        lib.ref_p2c_batch.argtypes = [C.c_int, _dp, C.c_double, _dp]
        lib.ref_hllc_batch.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]
        lib.ref_compute_dt.restype = C.c_double
        lib.ref_compute_dt.argtypes = [_dp, C.c_int, C.c_int, C.c_int] + [C.c_double] * 5
        lib.refg_create.restype = C.c_void_p
        lib.refg_create.argtypes = [C.c_int] * 3 + [C.c_double] * 7
        lib.refg_destroy.argtypes = [C.c_void_p]
        lib.refg_set_params.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
        lib.refg_set_bc.argtypes = [C.c_void_p, C.c_int]
        lib.refg_set_phi.argtypes = [C.c_void_p, _dp]
        lib.refg_solve.argtypes = [C.c_void_p, _dp]
        lib.refg_get.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        lib.refg_diag.argtypes = [C.c_void_p, _ip, _dp]
        lib.refg_moments.argtypes = [C.c_void_p, _dp, _dp]
        lib.refg_boundary.argtypes = [C.c_void_p, _dp, _dp]
        _rlib = lib
    return _rlib


class Grid:
    """Grid description shared by the oracle, the reference and the device."""

    def __init__(self, nx, ny, nz, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), G=0.0,
                 gamma=1.4, cfl=0.4, bc=(2, 2, 2, 2, 2, 2)):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.box = tuple(float(b) for b in box)
        self.G, self.gamma, self.cfl = float(G), float(gamma), float(cfl)
        self.bc = tuple(int(b) for b in bc)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz


class OracleSolver:
    """The oracle's solver with the reference semantics."""

    def __init__(self, grid, omega=1.5, max_iter=10000, tol=1e-6, lmax=2,
                 moment_mode=MOM_SEQUENTIAL, hoist_bc=1, grav_bc=GRAV_ISOLATED):
        self.lib = oracle_lib()
        self.g = grid
        self.h = self.lib.or_create(grid.nx, grid.ny, grid.nz, *grid.box, grid.G)
        bc = np.array(grid.bc, dtype=np.int32)
        self.lib.or_set_hydro(self.h, grid.gamma, grid.cfl, iptr(bc))
        self.lib.or_set_gravity(self.h, omega, max_iter, tol, lmax, moment_mode,
                                hoist_bc, grav_bc)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        assert U.size == 5 * self.g.ncells
        self.lib.or_set_state(self.h, dptr(U))

    def state(self):
        U = np.empty(5 * self.g.ncells)
        self.lib.or_get_state(self.h, dptr(U))
        return U.reshape(self.g.nx, self.g.ny, self.g.nz, 5)

    def set_phi(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        self.lib.or_set_phi(self.h, dptr(phi))

    def phi(self):
        p = np.empty(self.g.ncells)
        self.lib.or_get_phi(self.h, dptr(p))
        return p.reshape(self.g.shape)

    def gfield(self):
        a = [np.empty(self.g.ncells) for _ in range(3)]
        self.lib.or_get_g(self.h, *[dptr(x) for x in a])
        return [x.reshape(self.g.shape) for x in a]

    def grav_g(self):
        """The gravity solver's own g (written by a converged solve)."""
        a = [np.empty(self.g.ncells) for _ in range(3)]
        self.lib.or_get_grav_g(self.h, *[dptr(x) for x in a])
        return [x.reshape(self.g.shape) for x in a]

    @property
    def dt(self):
        return self.lib.or_get_dt(self.h)

    def diag(self):
        it = np.zeros(1, dtype=np.int32)
        st = np.zeros(1, dtype=np.int32)
        res = np.zeros(1)
        self.lib.or_get_diag(self.h, iptr(it), dptr(res), iptr(st))
        return int(it[0]), float(res[0]), int(st[0])

    def step(self):
        return self.lib.or_step(self.h)

    def run(self, t_end, max_steps=0, log_len=100000):
        t = np.zeros(1)
        dts = np.zeros(log_len)
        its = np.zeros(log_len, dtype=np.int32)
        n = self.lib.or_run(self.h, t_end, max_steps, dptr(t), dptr(dts), iptr(its))
        return n, float(t[0]), dts[:n].copy(), its[:n].copy()

    def gravity_solve(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        return self.lib.or_gravity_solve(self.h, dptr(rho))

    def moments(self, rho, lmax_hi=0):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        out = np.zeros(13)
        hi = np.zeros(256)
        self.lib.or_moments(self.h, dptr(rho), dptr(out), dptr(hi))
        return out, hi

    def boundary_phi(self, rho, phi):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1).copy()
        self.lib.or_boundary_phi(self.h, dptr(rho), dptr(phi))
        return phi.reshape(self.g.shape)

    def rel_residual(self, rho, phi):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        return self.lib.or_rel_residual(self.h, dptr(rho), dptr(phi))


class RefGravity:
    """The compiled reference's gravity solver on its own (gravity_solver_3d.c)."""

    def __init__(self, grid, omega=1.5, max_iter=10000, tol=1e-6, grav_bc=GRAV_ISOLATED):
        self.lib = ref_lib()
        assert self.lib is not None
        self.g = grid
        self.h = self.lib.refg_create(grid.nx, grid.ny, grid.nz, *grid.box, grid.G)
        self.lib.refg_set_params(self.h, omega, max_iter, tol)
        self.lib.refg_set_bc(self.h, grav_bc)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.refg_destroy(self.h)
            self.h = None

    def set_phi(self, phi):
        self._phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1).copy()
        self.lib.refg_set_phi(self.h, dptr(self._phi))

    def solve(self, rho):
        """gravity_solve: the potential, then g when it converged.  The
        solver keeps a pointer to rho, so it is held here."""
        self._rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1).copy()
        return self.lib.refg_solve(self.h, dptr(self._rho))

    def fields(self):
        a = [np.empty(self.g.ncells) for _ in range(4)]
        self.lib.refg_get(self.h, *[dptr(x) for x in a])
        return [x.reshape(self.g.shape) for x in a]

    def diag(self):
        it = np.zeros(1, dtype=np.int32)
        res = np.zeros(1)
        self.lib.refg_diag(self.h, iptr(it), dptr(res))
        return int(it[0]), float(res[0])


class RefSolver:
    """The compiled reference's hydro solver (gravity created when G > 0)."""

    def __init__(self, grid, omega=1.5, max_iter=10000, tol=1e-6):
        self.lib = ref_lib()
        assert self.lib is not None
        self.g = grid
        self.h = self.lib.ref_create(grid.nx, grid.ny, grid.nz, *grid.box, grid.G)
        bc = np.array(grid.bc, dtype=np.int32)
        self.lib.ref_set_hydro(self.h, grid.gamma, grid.cfl, iptr(bc))
        self.lib.ref_set_grav(self.h, omega, max_iter, tol)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_destroy(self.h)
            self.h = None

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        self.lib.ref_set_state(self.h, dptr(U))

    def state(self):
        U = np.empty(5 * self.g.ncells)
        self.lib.ref_get_state(self.h, dptr(U))
        return U.reshape(self.g.nx, self.g.ny, self.g.nz, 5)

    def phi(self):
        p = np.zeros(self.g.ncells)
        self.lib.ref_get_phi(self.h, dptr(p))
        return p.reshape(self.g.shape)

    def gfield(self):
        a = [np.empty(self.g.ncells) for _ in range(3)]
        self.lib.ref_get_g(self.h, *[dptr(x) for x in a])
        return [x.reshape(self.g.shape) for x in a]

    def diag(self):
        it = np.zeros(1, dtype=np.int32)
        res = np.zeros(1)
        self.lib.ref_diag(self.h, iptr(it), dptr(res))
        return int(it[0]), float(res[0])

    @property
    def dt(self):
        return self.lib.ref_dt(self.h)

    def step(self):
        self.lib.ref_step(self.h)

    def run(self, t_end, max_steps=0, log_len=100000):
        t = np.zeros(1)
        dts = np.zeros(log_len)
        its = np.zeros(log_len, dtype=np.int32)
        n = self.lib.ref_run(self.h, t_end, max_steps, dptr(t), dptr(dts), iptr(its))
        return n, float(t[0]), dts[:n].copy(), its[:n].copy()


def cell_centres(grid):
    x0, x1, y0, y1, z0, z1 = grid.box
    dx = (x1 - x0) / grid.nx
    dy = (y1 - y0) / grid.ny
    dz = (z1 - z0) / grid.nz
    x = x0 + (np.arange(grid.nx) + 0.5) * dx
    y = y0 + (np.arange(grid.ny) + 0.5) * dy
    z = z0 + (np.arange(grid.nz) + 0.5) * dz
    return np.meshgrid(x, y, z, indexing="ij")
