"""Device numerics that the parity tests rely on, and the FP32
defect-correction Poisson mode (no reference counterpart) against FP64."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu


def blob(shape, c=(0.5, 0.48, 0.52), R=0.3):
    X, Y, Z = ics.centres(shape)
    r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    return np.where(r < R, (1.0 - r / R) ** 2, 0.0) + 1e-4


@pytest.fixture(scope="module")
def tight():
    """A 32^3 potential converged to a 1e-11 relative residual in FP64."""
    shape = (32, 32, 32)
    rho = blob(shape)
    g = H.Gravity(*shape, G=1.0)
    g.set_lmax(4)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_params(1.5, 200000, 1e-11)
    assert g.solve_potential(rho) == 0
    return rho, g.phi()


def test_relative_residual_matches_oracle(tight):
    rho, phi = tight
    shape = rho.shape
    o = OracleSolver(Grid(*shape, G=1.0))
    g = H.Gravity(*shape, G=1.0)
    g.set_phi(phi)
    assert g.rel_residual(rho) == pytest.approx(o.rel_residual(rho, phi), rel=1e-9)
    assert o.rel_residual(rho, phi) <= 1e-11


@pytest.mark.parametrize("mode", [H.POISSON_FP64_SOR, H.POISSON_FP32_DC])
def test_poisson_modes_reach_tolerance_and_agree(tight, mode):
    rho, ref = tight
    shape = rho.shape
    g = H.Gravity(*shape, G=1.0)
    g.set_lmax(4)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_mode(mode)
    g.set_params(1.5, 200000, 1e-8)
    st = g.solve_potential(rho)
    assert st == 0
    phi = g.phi()
    o = OracleSolver(Grid(*shape, G=1.0))
    assert o.rel_residual(rho, phi) <= 1e-8  # residual in FP64 on the CPU
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    print(f"mode {mode}: rel L2 vs tight FP64 = {err:.3e}, sweeps {g.diagnostics()[0]}")
    assert err < 1e-7
