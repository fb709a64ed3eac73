"""numpy statement of the multigrid defect-correction Poisson solve, for the
tests (the device's DC_INNER_MULTIGRID mode; no reference counterpart).

Grid: the solver's nx x ny x nz cells with the outer layer as Dirichlet
boundary; the fine operator is gravity_solver_3d.c's 7-point Laplacian
(gravity_solver_3d.c:642-708: lap - factor * phi).  Everything on the
coarse levels is derived here from point positions alone, independently of
the device's closed-form coefficients:

- level l >= 1 keeps fine cells 0 and c 2^l - 1 (c >= 1), both end points
  on the fine boundary; halving stops when an interval count is odd or a
  direction has fewer than 5 points;
- operators: the three-point second difference on the points' positions,
  summed over directions;
- prolongation: linear interpolation in position, interior points only;
- restriction: prolongation's transpose weighted by dual-cell lengths,
  R = diag(1 / V_coarse) P^T diag(V_fine);
- smoothing: red-black SOR (red = even i + j + k first), the coarsest
  level max(n) Gauss-Seidel sweep pairs.

The outer loop: FP64 residual, inner V-cycles in `dtype`, phi += e, stop on
||r|| / ||f|| <= tol, V-cycles per round sized from the last round's
contraction to aim at a 1e-4 reduction (1 .. 16).
"""
import math

import numpy as np


def level_positions(shape):
    """[(X, Y, Z)] per level, positions in fine cells."""
    out = [tuple(np.arange(n, dtype=float) for n in shape)]
    c, l = [n // 2 + 1 for n in shape], 1
    while True:
        out.append(tuple(np.array([0.0] + [k * 2.0 ** l - 1.0 for k in range(1, m)])
                         for m in c))
        if not (all((m - 1) % 2 == 0 for m in c) and min(c) >= 5):
            return out
        c, l = [(m - 1) // 2 + 1 for m in c], l + 1


def op_1d(X, h):
    """Second-difference matrix on positions X (fine cells of length h)."""
    n = len(X)
    A = np.zeros((n, n))
    for i in range(1, n - 1):
        hl, hr = (X[i] - X[i - 1]) * h, (X[i + 1] - X[i]) * h
        s = 2.0 / (hl + hr)
        A[i, i - 1], A[i, i + 1], A[i, i] = s / hl, s / hr, -(s / hl + s / hr)
    return A


def interp_1d(Xf, Xc):
    P = np.zeros((len(Xf), len(Xc)))
    for i in range(1, len(Xf) - 1):
        x = Xf[i]
        c = int(np.searchsorted(Xc, x))
        if Xc[c] == x:
            P[i, c] = 1.0
        else:
            x0, x1 = Xc[c - 1], Xc[c]
            P[i, c - 1], P[i, c] = (x1 - x) / (x1 - x0), (x - x0) / (x1 - x0)
    return P


def dual_lengths(X):
    V = np.zeros(len(X))
    V[1:-1] = (X[2:] - X[:-2]) / 2.0
    return V


def restrict_1d(Xf, Xc):
    P, Vf, Vc = interp_1d(Xf, Xc), dual_lengths(Xf), dual_lengths(Xc)
    R = np.zeros((len(Xc), len(Xf)))
    for c in range(1, len(Xc) - 1):
        R[c] = P[:, c] * Vf / Vc[c]
    return R


def along_axes(mats, v):
    """Apply one matrix along each axis in turn (x, then y, then z)."""
    for ax, m in enumerate(mats):
        v = np.moveaxis(np.tensordot(m, v, axes=([1], [ax])), 0, ax)
    return v


def colour_masks(shape):
    i, j, k = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    par = (i + j + k) & 1
    return [inner & (par == 0), inner & (par == 1)]


class Multigrid:
    def __init__(self, shape, h, nu1=2, nu2=2, omega=1.0, dtype=np.float32):
        """shape: fine cells; h: (dx, dy, dz)."""
        self.shape, self.h, self.dtype = tuple(shape), tuple(h), dtype
        self.nu1, self.nu2, self.omega = nu1, nu2, omega
        self.pos = level_positions(shape)
        self.A = [[op_1d(X[a], h[a]).astype(dtype) for a in range(3)] for X in self.pos]
        self.P = [[interp_1d(self.pos[l][a], self.pos[l + 1][a]).astype(dtype) for a in range(3)]
                  for l in range(len(self.pos) - 1)]
        self.R = [[restrict_1d(self.pos[l][a], self.pos[l + 1][a]).astype(dtype)
                   for a in range(3)] for l in range(len(self.pos) - 1)]
        self.masks = [colour_masks(tuple(len(x) for x in X)) for X in self.pos]
        self.diag = []
        for A in self.A:
            d = [np.diag(m) for m in A]
            self.diag.append(d[0][:, None, None] + d[1][None, :, None] + d[2][None, None, :])

    def apply(self, l, u):
        """The level-l operator on u (0 on the boundary)."""
        out = np.zeros_like(u)
        for ax, m in enumerate(self.A[l]):
            out += np.moveaxis(np.tensordot(m, u, axes=([1], [ax])), 0, ax)
        inner = self.masks[l][0] | self.masks[l][1]
        return np.where(inner, out, 0).astype(u.dtype)

    def sweep(self, l, u, f, omega):
        D = self.diag[l]
        om = self.dtype(omega)
        for m in self.masks[l]:
            off = self.apply(l, u) - D * u
            gs = (f - off) / np.where(m, D, 1)
            u[m] = u[m] + om * (gs[m] - u[m])

    def cycle(self, l, u, f):
        """V-cycle from level l on A_l u = f (u updated in place)."""
        if l == len(self.pos) - 1:
            for _ in range(max(u.shape)):
                self.sweep(l, u, f, 1.0)
            return
        for _ in range(self.nu1):
            self.sweep(l, u, f, self.omega)
        q = self.residual(l, u, f)
        fc = along_axes(self.R[l], q).astype(self.dtype)
        uc = np.zeros_like(fc)
        self.cycle(l + 1, uc, fc)
        u += along_axes(self.P[l], uc).astype(self.dtype)
        for _ in range(self.nu2):
            self.sweep(l, u, f, self.omega)

    def residual(self, l, u, f):
        q = (f - self.apply(l, u)).astype(u.dtype)
        return np.where(self.masks[l][0] | self.masks[l][1], q, 0).astype(u.dtype)


def dc_multigrid(phi, f, h, tol, max_cycles=1000, **kw):
    """Defect correction with multigrid inner solves; phi (FP64, boundary
    values set) is updated in place.  Returns (V-cycles, relative residual,
    relative residual after each round)."""
    mg = Multigrid(phi.shape, h, **kw)
    mg64 = Multigrid(phi.shape, h, dtype=np.float64)
    inner = mg.masks[0][0] | mg.masks[0][1]

    def residual():
        r = mg64.residual(0, phi, f)
        return r, math.sqrt(float(np.sum(r * r)) / float(np.sum(np.where(inner, f, 0.0) ** 2)))

    r, rel = residual()
    hist = [rel]
    cycles, k = 0, 3
    for _ in range(64):
        if rel <= tol or cycles >= max_cycles:
            break
        nk = min(k, max_cycles - cycles)
        e = np.zeros(phi.shape, dtype=mg.dtype)
        r32 = r.astype(mg.dtype)
        for _ in range(nk):
            mg.cycle(0, e, r32)
        cycles += nk
        phi += e.astype(np.float64)
        r, rnew = residual()
        hist.append(rnew)
        red = rnew / rel
        if 0.0 < red < 1.0:
            rate = red ** (1.0 / nk)
            want = max(1e-4, 0.5 * tol / rnew)
            k = max(1, min(int(math.ceil(math.log(want) / math.log(rate))), 16))
        else:
            k = min(2 * k, 16)
        rel = rnew
    return cycles, rel, hist
