"""The device with MOMENT_ORDER_SEQUENTIAL against the outputs of the
compiled reference itself (tests/golden/gravity.npz and selfgrav.npz, made
by tests/golden/make_golden.py from gravity_solver_3d.c and
hydro_solver_3d.c): the 13 moments, the multipole boundary potential, full
solves at tolerance 1e-6 and 1e-8 (iteration count, residual, potential
and acceleration), and ten self-gravity steps of the reference's run loop,
all bit for bit."""
import os

import numpy as np
import pytest

import hydrocode_astro1_amd as H

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLDEN, name))


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def test_moments_boundary_and_solves_equal_the_reference():
    f = gold("gravity.npz")
    rho = np.ascontiguousarray(f["rho"])
    print("reference gravity fixture, sequential moment order", flush=True)
    g = H.Gravity(16, 16, 16, G=1.0)
    g.set_moment_order(H.MOMENT_ORDER_SEQUENTIAL)
    mom, _ = g.moments(rho)
    assert same_bits(mom, f["moments"])
    bphi = g.boundary_potential(rho)
    shell = np.ones((16, 16, 16), bool)
    shell[1:-1, 1:-1, 1:-1] = False
    assert same_bits(bphi[shell], f["boundary_phi"].reshape(16, 16, 16)[shell])
    g.close()
    for tag, tol in (("6", 1e-6), ("8", 1e-8)):
        g = H.Gravity(16, 16, 16, G=1.0)
        g.set_moment_order(H.MOMENT_ORDER_SEQUENTIAL)
        g.set_params(1.5, 10000, tol)
        assert g.solve(rho) == int(f[f"status_{tag}"])
        it, res = g.diagnostics()
        assert it == int(f[f"iters_{tag}"])
        assert res == float(f[f"res_{tag}"])
        assert same_bits(g.phi(), f[f"phi_{tag}"].reshape(16, 16, 16))
        for a, c in zip(g.g(), ("gx", "gy", "gz")):
            assert same_bits(a, f[f"{c}_{tag}"].reshape(16, 16, 16))
        g.close()


def test_selfgravity_run_equals_the_reference():
    f = gold("selfgrav.npz")
    print("reference self-gravity fixture, sequential moment order", flush=True)
    d = H.Solver(20, 20, 20, G=1.0, gamma=2.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
    d.gravity.set_params(1.5, 10000, 1e-8)
    d.gravity.set_moment_order(H.MOMENT_ORDER_SEQUENTIAL)
    d.set_state(np.ascontiguousarray(f["U0"]))
    n, t, dts, its = d.run(1e9, max_steps=10)
    assert n == 10
    assert list(its) == list(f["its"])
    assert same_bits(dts, f["dts"])
    assert same_bits(d.state(), f["U"])
    d.close()
