"""Two SOR iterations per pass (sor2) against one-iteration sweeps on the
device and against the oracle: potential bit for bit, the same iteration
counts and status under both stop rules, convergence landing on either
iteration of a pass, iteration caps of every parity, grids that are not
multiples of the tile, and the FP32 correction sweeps."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics
from oracle_lib import MOM_DDTREE, Grid, OracleSolver

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def field(shape, seed=0):
    X, Y, Z = ics.centres(shape)
    rng = np.random.default_rng(seed)
    r = np.sqrt((X - 0.47) ** 2 + (Y - 0.52) ** 2 + (Z - 0.5) ** 2)
    return np.where(r < 0.33, 1.0 - r / 0.33, 0.0) + 1e-6 + 0.02 * rng.random(shape)


def solver(monkeypatch, fused, shape, tol, max_iter=20000, crit=H.SOR_CRIT_MAX_UPDATE,
           mode=H.POISSON_FP64_SOR, omega=1.5):
    monkeypatch.setenv("HYDRO_MI355X_SOR_FUSED", "1" if fused else "0")
    g = H.Gravity(*shape, G=1.0)
    g.set_params(omega, max_iter, tol)
    g.set_criterion(crit)
    g.set_mode(mode)
    return g


def run_pair(monkeypatch, shape, rho, tol, **kw):
    out = []
    for fused in (False, True):
        g = solver(monkeypatch, fused, shape, tol, **kw)
        st = g.solve_potential(rho)
        it, res = g.diagnostics()
        out.append((st, it, res, g.phi()))
    return out


def test_fused_equals_single_max_update(monkeypatch):
    # grids on and off the tile size, and a spread of tolerances so that
    # convergence falls on both the first and the second iteration of a pass
    parities = set()
    for shape in [(40, 36, 70), (19, 23, 17), (70, 18, 40)]:
        rho = field(shape)
        for tol in (1e-5, 3e-6, 1e-6, 4e-7, 1e-7, 5e-8):
            (s0, i0, r0, p0), (s1, i1, r1, p1) = run_pair(monkeypatch, shape, rho, tol)
            assert (s1, i1) == (s0, i0), (shape, tol)
            assert r1 == r0, (shape, tol)
            assert same_bits(p1, p0), (shape, tol)
            parities.add(i0 % 2)
    assert parities == {0, 1}


@pytest.mark.parametrize("max_iter", [1, 2, 3, 4, 5, 8, 9])
def test_fused_iteration_caps(monkeypatch, max_iter):
    shape = (21, 26, 34)
    rho = field(shape, 1)
    (s0, i0, r0, p0), (s1, i1, r1, p1) = run_pair(monkeypatch, shape, rho, 1e-12,
                                                  max_iter=max_iter)
    assert (s0, i0) == (-1, max_iter)
    assert (s1, i1, r1) == (s0, i0, r0)
    assert same_bits(p1, p0)


def test_fused_against_oracle_warm_starts(monkeypatch):
    shape = (34, 30, 38)
    rho = field(shape, 2)
    o = OracleSolver(Grid(*shape, G=1.0), tol=1e-7, moment_mode=MOM_DDTREE)
    g = solver(monkeypatch, True, shape, 1e-7)
    for rep in range(3):
        r = rho * (1.0 + 0.02 * rep)
        assert o.gravity_solve(r) == g.solve(r) == 0
        assert g.diagnostics()[0] == o.diag()[0]
        assert same_bits(g.phi(), o.phi())


def test_fused_relative_residual_mode(monkeypatch):
    shape = (48, 40, 44)
    rho = field(shape, 3)
    (s0, i0, r0, p0), (s1, i1, r1, p1) = run_pair(monkeypatch, shape, rho, 1e-9,
                                                  crit=H.SOR_CRIT_REL_RESID)
    assert s0 == s1 == 0 and i0 == i1
    assert same_bits(p1, p0)


def test_fused_fp32_correction(monkeypatch):
    shape = (48, 40, 44)
    rho = field(shape, 4)
    (s0, i0, r0, p0), (s1, i1, r1, p1) = run_pair(monkeypatch, shape, rho, 1e-9,
                                                  crit=H.SOR_CRIT_REL_RESID,
                                                  mode=H.POISSON_FP32_DC)
    assert s0 == s1 == 0 and i0 == i1
    assert same_bits(p1, p0)
