"""The multigrid inner solve of the FP32 defect correction (DC_INNER_MULTIGRID,
no reference counterpart) on the device: on a zero-Dirichlet problem it
lands on the exact discrete solution (sine transform) and takes the V-cycle
count of the numpy statement (tests/mg_reference.py); with the isolated
multipole boundary it reaches the tolerance and agrees with FP64 SOR."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import nn
import mg_reference as M
from test_gpu_numerics import blob, tight  # noqa: F401  (module fixture)
from test_mg_reference import dirichlet_solve
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu

# the device's default smoothing: V(1,1) with SOR factor 1.15
SMOOTHING = dict(nu1=1, nu2=1, omega=1.15)


@pytest.mark.parametrize("shape,box", [((32, 32, 32), (1.0, 1.0, 1.0)),
                                       ((48, 32, 40), (1.2, 1.0, 0.9))])
def test_dirichlet_exact_solution_and_cycle_count(shape, box):
    print(f"multigrid Dirichlet {shape} box {box}", flush=True)
    G = 1.0
    g = H.Gravity(*shape, G=G, box=(0.0, box[0], 0.0, box[1], 0.0, box[2]))
    g.set_boundary(H.GRAV_BC_DIRICHLET)
    g.set_mode(H.POISSON_FP32_DC)
    g.set_dc_inner(H.DC_INNER_MULTIGRID)
    g.set_params(1.5, 1000, 1e-10)
    rho = blob(shape)
    assert g.solve_potential(rho) == 0
    phi = g.phi()
    cycles = g.diagnostics()[0]
    h = tuple(b / n for b, n in zip(box, shape))
    f = np.zeros(shape)
    f[1:-1, 1:-1, 1:-1] = 4.0 * np.pi * G * rho[1:-1, 1:-1, 1:-1]
    exact = dirichlet_solve(f, h)
    err = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    ref = np.zeros(shape)
    ref_cycles, ref_rel, hist = M.dc_multigrid(ref, f, h, tol=1e-10, **SMOOTHING)
    print(f"  device {cycles} V-cycles, rel L2 vs exact {err:.2e}; numpy {ref_cycles} {hist}")
    assert g.s.last_rel_residual <= 1e-10
    assert err < 1e-9
    assert abs(cycles - ref_cycles) <= 1


def test_multigrid_rejects_odd_grids():
    g = H.Gravity(33, 32, 32, G=1.0)
    with pytest.raises(ValueError):
        g.set_dc_inner(H.DC_INNER_MULTIGRID)


@pytest.mark.parametrize("nn_mode", [False, True])
def test_isolated_boundary_reaches_tolerance_and_agrees(tight, nn_mode):
    """32^3 blob with the l = 4 multipole boundary: the 1e-8 solve agrees
    with a 1e-11 FP64 SOR potential to 1e-7 relative L2 (the FP64 residual
    is checked on the CPU by the oracle)."""
    rho, ref = tight
    shape = rho.shape
    g = H.Gravity(*shape, G=1.0)
    g.set_lmax(4)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_mode(H.POISSON_NN_FP32_DC if nn_mode else H.POISSON_FP32_DC)
    if nn_mode:
        g.set_nn_model(nn.Model(path=nn.DEFAULT_WEIGHTS))
    g.set_dc_inner(H.DC_INNER_MULTIGRID)
    g.set_params(1.5, 1000, 1e-8)
    assert g.solve_potential(rho) == 0
    phi = g.phi()
    o = OracleSolver(Grid(*shape, G=1.0))
    assert o.rel_residual(rho, phi) <= 1.05e-8
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    print(f"multigrid isolated: {g.diagnostics()[0]} V-cycles, rel L2 vs tight FP64 {err:.2e}, "
          f"guess {g.s.last_guess}")
    if nn_mode:
        assert g.s.last_guess in (0, 1)
    assert err < 1e-7
    assert g.diagnostics()[0] <= 12


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 48, 40)])
def test_lds_tail_gives_the_same_bits(shape, monkeypatch):
    """The coarse levels run in one workgroup in LDS (the default) or one
    phase per launch (HYDRO_MI355X_MG_TAIL=0): the same point operations in
    the same phase order, so the potential and the V-cycle count agree bit
    for bit.  32^3: the tail is every coarse level; 64 x 48 x 40: the last
    two of three."""
    rho = blob(shape)
    out = []
    for tail in ("1", "0"):
        print(f"multigrid tail={tail} {shape}", flush=True)
        monkeypatch.setenv("HYDRO_MI355X_MG_TAIL", tail)
        g = H.Gravity(*shape, G=1.0)
        g.set_lmax(4)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        g.set_mode(H.POISSON_FP32_DC)
        g.set_dc_inner(H.DC_INNER_MULTIGRID)
        g.set_params(1.5, 1000, 1e-9)
        assert g.solve_potential(rho) == 0
        out.append((g.diagnostics()[0], g.phi()))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
