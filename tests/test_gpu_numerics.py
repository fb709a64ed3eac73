"""Device numerics that the parity tests rely on, and the FP32
defect-correction Poisson mode (no reference counterpart) against FP64."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu


def test_reciprocal_division_is_ieee_division():
    """The reciprocal-corrected division (RN(1/b) and one Markstein
    correction, with an IEEE fallback; used by the step kernel when built
    with HYD_FASTDIV=1) must equal a / b in every bit, including the zero,
    tiny, huge and all-ones-significand cases that take the fallback."""
    rng = np.random.default_rng(99)
    n = 4_000_000
    a = rng.standard_normal(n) * 10.0 ** rng.uniform(-300, 300, n)
    b = np.abs(rng.standard_normal(n)) * 10.0 ** rng.uniform(-300, 300, n) + 1e-300
    # densities and momenta as the solver sees them
    a[:1_000_000] = rng.standard_normal(1_000_000) * 10.0 ** rng.uniform(-12, 4, 1_000_000)
    b[:1_000_000] = 10.0 ** rng.uniform(-5, 2, 1_000_000)
    # all-ones significands, exact powers of two, signed zeros, subnormals
    ones = np.frombuffer(np.uint64(0x3FFFFFFFFFFFFFFF).tobytes() * 1000, dtype=np.float64)
    b[1_000_000:1_001_000] = ones * 2.0 ** rng.integers(-60, 60, 1000)
    b[1_001_000:1_002_000] = 2.0 ** rng.integers(-900, 900, 1000)
    a[1_002_000:1_002_500] = 0.0
    a[1_002_500:1_003_000] = -0.0
    a[1_003_000:1_004_000] = 5e-324 * rng.integers(1, 1 << 40, 1000)
    b[1_004_000:1_005_000] = 1e-5  # the density floor
    assert H.lib().hydro_check_division(a.ctypes.data_as(H._dp), b.ctypes.data_as(H._dp), n) == 0


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
    assert o.rel_residual(rho, phi) <= 1.05e-8  # residual in FP64 on the CPU
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    print(f"mode {mode}: rel L2 vs tight FP64 = {err:.3e}, sweeps {g.diagnostics()[0]}")
    assert err < 1e-6
