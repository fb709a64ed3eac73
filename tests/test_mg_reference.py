"""The numpy multigrid defect-correction solve (tests/mg_reference.py) pinned
on the CPU, and the device's closed-form coarse-level coefficients checked
against it: restriction keeps constants, prolongation keeps linear
functions, a V-cycle contracts the residual the way multigrid must, and the
defect-correction loop lands on the exact discrete solution given by a
sine transform (DST-I diagonalises the Dirichlet 7-point Laplacian)."""
import numpy as np
import pytest
from scipy.fft import dstn, idstn

import mg_reference as M


def dirichlet_solve(f, h):
    """Exact solution of the 7-point system lap(u) - factor u = f on the
    interior, u = 0 on the boundary, by the type-I sine transform."""
    inner = f[1:-1, 1:-1, 1:-1]
    lam = []
    for m, hh in zip(inner.shape, h):
        k = np.arange(1, m + 1)
        lam.append(-4.0 / hh ** 2 * np.sin(np.pi * k / (2.0 * (m + 1))) ** 2)
    L = lam[0][:, None, None] + lam[1][None, :, None] + lam[2][None, None, :]
    u = np.zeros_like(f)
    u[1:-1, 1:-1, 1:-1] = idstn(dstn(inner, type=1) / L, type=1)
    return u


def interior_random(shape, rng):
    q = np.zeros(shape)
    q[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(s - 2 for s in shape))
    return q


def test_levels_of_a_256_grid():
    pos = M.level_positions((256, 256, 256))
    assert [len(X[0]) for X in pos] == [256, 129, 65, 33, 17, 9, 5, 3]
    assert list(pos[7][0]) == [0.0, 127.0, 255.0]   # end points on the fine boundary


def test_device_coefficients_match_positions():
    """The closed forms the device uses for level l (H = 2^l fine cells):
    a = 1/(H h)^2; at index 1 with hl = (H-1) h, hr = H h: am1 = 2/((hl+hr) hl),
    ap1 = 2/((hl+hr) hr); pw1 = (H-1)/(2H-1); r1 = (Hf-1, 2Hf, Hf)/(4Hf-1),
    Hf = H/2 -- against the matrices derived from positions."""
    h = 0.37
    pos = M.level_positions((64, 64, 64))
    for l in range(1, len(pos)):
        H, Hf = 2.0 ** l, 2.0 ** (l - 1)
        X, Xf = pos[l][0], pos[l - 1][0]
        A = M.op_1d(X, h)
        hl, hr = (H - 1) * h, H * h
        assert A[1, 0] == pytest.approx(2 / ((hl + hr) * hl), rel=1e-14)
        assert A[1, 2] == pytest.approx(2 / ((hl + hr) * hr), rel=1e-14)
        if len(X) > 3:  # index 2 is interior
            assert A[2, 1] == pytest.approx(1 / (H * h) ** 2, rel=1e-14)
            assert A[2, 3] == pytest.approx(1 / (H * h) ** 2, rel=1e-14)
        R = M.restrict_1d(Xf, X)
        # coarse point 1 sits on finer point 2 (level 1: fine cell 1, i.e. index 1)
        centre = 1 if l == 1 else 2
        np.testing.assert_allclose(R[1, centre - 1:centre + 2],
                                   np.array([Hf - 1, 2 * Hf, Hf]) / (4 * Hf - 1), rtol=1e-14)
        if len(X) > 3:  # coarse point 2 sits on finer point 3 (level 1) or 4
            c2 = 3 if l == 1 else 4
            np.testing.assert_allclose(R[2, c2 - 1:c2 + 2], [0.25, 0.5, 0.25], rtol=1e-14)
        if l + 1 < len(pos):
            P = M.interp_1d(X, pos[l + 1][0])
            assert P[1, 1] == pytest.approx((H - 1) / (2 * H - 1), rel=1e-14)
            assert P[3, 1] == pytest.approx(0.5) and P[3, 2] == pytest.approx(0.5)


def test_restriction_keeps_constants_and_prolongation_keeps_linear():
    shape = (32, 24, 40)
    mg = M.Multigrid(shape, (1.0, 1.0, 1.0), dtype=np.float64)
    for l in range(len(mg.pos) - 1):
        fine = tuple(len(x) for x in mg.pos[l])
        ones = np.zeros(fine)
        ones[1:-1, 1:-1, 1:-1] = 1.0
        fc = M.along_axes(mg.R[l], ones)
        np.testing.assert_allclose(fc[1:-1, 1:-1, 1:-1], 1.0, rtol=1e-13)
        Xc, Xf = mg.pos[l + 1], mg.pos[l]
        uc = 0.3 * Xc[0][:, None, None] - 1.1 * Xc[1][None, :, None] + 0.7 * Xc[2][None, None, :]
        uf = M.along_axes(mg.P[l], uc)
        expect = 0.3 * Xf[0][:, None, None] - 1.1 * Xf[1][None, :, None] + 0.7 * Xf[2][None, None, :]
        np.testing.assert_allclose(uf[1:-1, 1:-1, 1:-1], expect[1:-1, 1:-1, 1:-1], atol=1e-11)


@pytest.mark.parametrize("n,h,bound", [(32, (1.0, 1.0, 1.0), 1 / 7), (48, (1.0, 1.0, 1.0), 1 / 7),
                                       (32, (1.0, 1.3, 0.8), 0.2)])
def test_vcycle_contraction(n, h, bound):
    """FP64 V(2,2): the residual of L e = r falls at least 7x per cycle on
    cubic cells (textbook red-black multigrid: about 0.1), 5x with cells
    whose squared aspect ratio is 2.6 (a point smoother loses some grip)."""
    rng = np.random.default_rng(n)
    r = interior_random((n, n, n), rng)
    e = np.zeros_like(r)
    mg = M.Multigrid(r.shape, h, dtype=np.float64)
    norms = [np.linalg.norm(r)]
    for _ in range(6):
        mg.cycle(0, e, r)
        norms.append(np.linalg.norm(mg.residual(0, e, r)))
    rates = np.array(norms[1:]) / np.array(norms[:-1])
    assert np.all(rates < bound), rates


def test_defect_correction_reaches_the_exact_discrete_solution():
    n = 32
    h = (1.0 / n, 1.2 / n, 0.8 / n)
    rng = np.random.default_rng(3)
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = np.exp(-((x - 0.4) ** 2 + (y - 0.55) ** 2 + (z - 0.5) ** 2) / 0.01)
    rho += 0.05 * rng.random(rho.shape)
    f = np.zeros_like(rho)
    f[1:-1, 1:-1, 1:-1] = 4.0 * np.pi * rho[1:-1, 1:-1, 1:-1]
    exact = dirichlet_solve(f, h)
    phi = np.zeros_like(f)
    cycles, rel, hist = M.dc_multigrid(phi, f, h, tol=1e-10)
    assert rel <= 1e-10
    assert cycles <= 16, (cycles, hist)
    err = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    assert err < 1e-9, err
