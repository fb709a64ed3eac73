"""The mixed-precision and learned-guess Poisson configurations at their own
sizes, to the bars they are held to: every potential reaches a relative
residual ||4 pi G rho - L_h phi|| / ||4 pi G rho|| <= 1e-8 as the CPU oracle
evaluates it in FP64 (or_rel_residual, interior cells), and lies within 1e-7
relative L2 of a tight FP64 potential.

 - Mixed precision: FP32 defect correction (FP64 residual and accumulation,
   FP32 red-black SOR sweeps or multigrid V-cycles on the correction) on the
   stellar-explosion density after 20 self-gravity steps at 256^3, warm-
   started from the previous step's potential, against plain FP64 SOR to a
   1e-11 residual; and on its 2x2x2 prolongation to 512^3 (multigrid inner
   solves), against the same method run on to a 1e-11 FP64 residual (near
   the floor of FP32 corrections, about 4e-12 here; plain FP64 SOR needs
   minutes at 512^3).  (SOR-swept FP32
   correction at 512^3 needs about 10^5 sweeps, half a minute; the bench's
   snapshot study times it.)
 - Learned guess: the committed network (trained on seeds 0..1999) on
   held-out 128^3 fields (seeds 2000..2003) and on the 256^3 explosion
   density, cold starts, with the same solver with and without the guess:
   the guess is taken and needs fewer SOR sweeps; multigrid needs no more
   V-cycles with it.  On a warm start the previous potential has the smaller
   residual and is kept.
Each test prints one JSON record of what it measured."""
import json

import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics, nn
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu

TOL = 1e-8
PHI_BAR = 1e-7


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def record(name, **kw):
    print("[record] " + json.dumps({"test": name, **kw}), flush=True)


def oracle_residual(n, rho, phi, G=1.0):
    o = OracleSolver(Grid(n, n, n, G=G, gamma=5.0 / 3.0))
    r = o.rel_residual(rho, phi)
    del o
    return r


def gravity(n, lmax=6):
    g = H.Gravity(n, n, n, G=1.0)
    g.set_lmax(lmax)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    return g


def solve(g, mode, inner, rho, phi0, tol=TOL, omega=1.5, max_iter=2000000):
    g.set_mode(mode)
    g.set_dc_inner(inner)
    g.set_params(omega, max_iter, tol)
    g.set_phi(phi0)
    st = g.solve_potential(rho)
    s = g.s
    info = {"status": st, "iterations": s.last_iterations, "ms": s.last_solve_ms,
            "device_rel_residual": s.last_rel_residual}
    if mode == H.POISSON_NN_FP32_DC:
        info.update(guess_used=s.last_guess == 1, guess_rel_residual=s.last_guess_rel_residual,
                    start_rel_residual=s.last_warm_rel_residual)
    return g.phi(), info


@pytest.fixture(scope="module")
def explosion_snapshot():
    """rho after 20 self-gravity steps of the 256^3 explosion (the
    reference's Poisson settings, l = 6) and the potential of step 20."""
    n = 256
    print(f"[configs] explosion {n}^3, 20 self-gravity steps", flush=True)
    s = H.Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
    s.gravity.set_lmax(6)
    s.gravity.set_params(1.5, 10000, 1e-6)
    U = ics.explosion_n3(n)
    rho0 = np.ascontiguousarray(U[..., 0])
    s.set_state(U)
    del U
    for _ in range(20):
        s.step()
    rho = np.ascontiguousarray(s.state()[..., 0])
    H.lib().gravity_sync_to_host(s.gravity.h)
    phi_prev = s.gravity.phi()
    s.close()
    return rho0, rho, phi_prev


@pytest.mark.parametrize("inner", ["sor", "multigrid"])
def test_mixed_precision_256(explosion_snapshot, inner):
    n = 256
    _, rho, phi_prev = explosion_snapshot
    g = gravity(n)
    IN = H.DC_INNER_SOR if inner == "sor" else H.DC_INNER_MULTIGRID
    print(f"[configs] FP32 defect correction ({inner}) at {n}^3", flush=True)
    phi, info = solve(g, H.POISSON_FP32_DC, IN, rho, phi_prev)
    tight, tinfo = solve(g, H.POISSON_FP64_SOR, H.DC_INNER_SOR, rho, phi, tol=1e-11)
    g.close()
    r = oracle_residual(n, rho, phi)
    err = rel_l2(phi, tight)
    record(f"mixed_precision_256_{inner}", rel_residual_oracle=r, rel_l2_vs_fp64=err,
           fp64_tight=tinfo, **info)
    assert info["status"] == 0 and tinfo["status"] == 0
    assert r <= TOL
    assert err <= PHI_BAR


def test_mixed_precision_512_multigrid(explosion_snapshot):
    n = 512
    _, rho256, phi256 = explosion_snapshot
    up = lambda a: np.repeat(np.repeat(np.repeat(a, 2, 0), 2, 1), 2, 2)  # noqa: E731
    rho, phi_prev = up(rho256), up(phi256)
    g = gravity(n)
    print(f"[configs] FP32 defect correction (multigrid) at {n}^3", flush=True)
    phi, info = solve(g, H.POISSON_FP32_DC, H.DC_INNER_MULTIGRID, rho, phi_prev)
    tight, tinfo = solve(g, H.POISSON_FP32_DC, H.DC_INNER_MULTIGRID, rho, phi, tol=1e-11)
    g.close()
    r = oracle_residual(n, rho, phi)
    rt = oracle_residual(n, rho, tight)
    err = rel_l2(phi, tight)
    record("mixed_precision_512_multigrid", rel_residual_oracle=r, tight_rel_residual_oracle=rt,
           rel_l2_vs_tight=err, tight=tinfo, **info)
    assert info["status"] == 0 and tinfo["status"] == 0
    assert r <= TOL and rt <= 1e-11
    assert err <= PHI_BAR


@pytest.fixture(scope="module")
def model():
    return nn.Model(path=nn.DEFAULT_WEIGHTS)


def guess_against_plain(g, model, rho, inner):
    zero = np.zeros_like(rho, dtype=np.float64)
    g.set_nn_model(None)
    phi_a, a = solve(g, H.POISSON_FP32_DC, inner, rho, zero, omega=1.9)
    g.set_nn_model(model)
    phi_b, b = solve(g, H.POISSON_NN_FP32_DC, inner, rho, zero, omega=1.9)
    g.set_nn_model(None)
    return phi_a, a, phi_b, b


@pytest.mark.parametrize("seed", [2000, 2001, 2002, 2003])
def test_learned_guess_on_held_out_fields(model, seed):
    n = 128
    ds = nn.Dataset(n, 1)
    gen = nn.make_solver(n)
    assert ds.fill(0, nn.field_params(seed), gen) == 0
    gen.close()
    rho32, _, _ = ds.get(0)
    ds.close()
    rho = rho32.astype(np.float64)
    g = gravity(n)
    out = {}
    for name, inner in (("sor", H.DC_INNER_SOR), ("multigrid", H.DC_INNER_MULTIGRID)):
        print(f"[configs] held-out field {seed}, {name} inner, with and without the guess",
              flush=True)
        phi_a, a, phi_b, b = guess_against_plain(g, model, rho, inner)
        ra, rb = oracle_residual(n, rho, phi_a), oracle_residual(n, rho, phi_b)
        out[name] = {"without": a, "with": b, "rel_residual_oracle": [ra, rb],
                     "rel_l2_between": rel_l2(phi_b, phi_a)}
        assert a["status"] == 0 and b["status"] == 0
        assert ra <= TOL and rb <= TOL
        assert rel_l2(phi_b, phi_a) <= PHI_BAR
        assert b["guess_used"]
        if name == "sor":
            assert b["iterations"] < a["iterations"]
        else:
            assert b["iterations"] <= a["iterations"]
    g.close()
    record(f"learned_guess_held_out_{seed}", **out)


def test_learned_guess_on_the_explosion_256(model, explosion_snapshot):
    """Applied at 256^3 on the explosion's initial density (the network runs
    on its 2x restriction): cold, the guess is taken and saves SOR sweeps;
    warm (the snapshot after 20 steps, from the previous potential) the
    previous potential is the better start and is kept."""
    n = 256
    rho0, rho, phi_prev = explosion_snapshot
    g = gravity(n)
    print("[configs] learned guess on the 256^3 explosion, cold start", flush=True)
    phi_a, a, phi_b, b = guess_against_plain(g, model, rho0, H.DC_INNER_SOR)
    ra, rb = oracle_residual(n, rho0, phi_a), oracle_residual(n, rho0, phi_b)
    g.set_nn_model(model)
    phi_w, w = solve(g, H.POISSON_NN_FP32_DC, H.DC_INNER_SOR, rho, phi_prev, omega=1.9)
    g.set_nn_model(None)
    g.close()
    record("learned_guess_explosion_256", cold_without=a, cold_with=b,
           rel_residual_oracle=[ra, rb], rel_l2_between=rel_l2(phi_b, phi_a), warm_with=w)
    assert a["status"] == 0 and b["status"] == 0 and w["status"] == 0
    assert ra <= TOL and rb <= TOL
    assert rel_l2(phi_b, phi_a) <= PHI_BAR
    assert b["guess_used"] and b["iterations"] < a["iterations"]
    assert not w["guess_used"]
