"""Device hydro step against the CPU oracle (bit-exact FP64).

The oracle restates hydro_solver_3d.c:162-811; these tests run the same
seeded inputs through the device library's C ABI and require identical bits
for every conserved field, the dt sequence and the step count."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu

BC_COMBOS = [(0,) * 6, (1,) * 6, (2,) * 6, (1, 2, 0, 0, 2, 1), (0, 1, 2, 2, 1, 0),
             (2, 1, 1, 2, 0, 0)]


def random_smooth(shape, gamma, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = ics.centres(shape)
    rho = 1 + 0.5 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.2 * rng.random(shape)
    u = 0.3 * np.sin(2 * np.pi * Z) + 0 * X
    v = rng.normal(0, 0.2, shape)
    w = 0.1 + 0 * X
    p = 1 + 0.3 * rng.random(shape)
    return ics.prim_to_cons(rho, u, v, w, p, gamma)


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def test_sod_tube_to_t02_bitwise():
    g = Grid(64, 16, 16, G=0.0, gamma=1.4, cfl=0.4, bc=(2,) * 6)
    U0 = ics.sod(64, 16, 16, 1.4)
    o = OracleSolver(g)
    o.set_state(U0)
    n_o, t_o, dts_o, _ = o.run(0.2)
    d = H.Solver(64, 16, 16, G=0.0, gamma=1.4, cfl=0.4, bc=(2,) * 6)
    d.set_state(U0)
    n_d, t_d, dts_d, _ = d.run(0.2)
    assert n_d == n_o
    assert t_d == t_o
    assert same_bits(dts_d, dts_o)
    assert same_bits(d.state(), o.state())


@pytest.mark.parametrize("bc", BC_COMBOS)
def test_random_state_each_bc_five_steps_bitwise(bc):
    shape = (12, 10, 14)
    g = Grid(*shape, G=0.0, gamma=5 / 3, cfl=0.3, bc=bc)
    U0 = random_smooth(shape, 5 / 3, seed=sum(bc) + 7)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*shape, G=0.0, gamma=5 / 3, cfl=0.3, bc=bc)
    d.set_state(U0)
    for _ in range(5):
        o.step()
        d.step()
        assert d.dt == o.dt
    assert same_bits(d.state(), o.state())


@pytest.mark.parametrize("cx", ["0", "5", "64"])
def test_tile_edges_bitwise(cx, monkeypatch):
    """A grid that is not a multiple of the 10 x 64 tile in j and k, with
    strong shocks and floors, and blocks of the automatic size, of 5 and of
    64 (tile, plane) pairs: runs that start and end inside a tile and runs
    that end one tile (nx = 70) and start the next."""
    monkeypatch.setenv("HYDRO_MI355X_CX", cx)
    shape = (70, 19, 131)
    rng = np.random.default_rng(3)
    X, Y, Z = ics.centres(shape)
    rho = np.where((X - 0.4) ** 2 + (Y - 0.5) ** 2 + (Z - 0.6) ** 2 < 0.05, 2.0, 1e-6) + 0 * X
    p = np.where(rho > 1, 50.0, 1e-8) + 0 * X
    u = rng.normal(0, 1, shape)
    U0 = ics.prim_to_cons(rho, u, 0 * X, -u, p, 1.4)
    bc = (2, 1, 0, 0, 1, 2)
    g = Grid(*shape, G=0.0, gamma=1.4, cfl=0.4, bc=bc)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*shape, G=0.0, gamma=1.4, cfl=0.4, bc=bc)
    d.set_state(U0)
    for _ in range(3):
        o.step()
        d.step()
    assert d.dt == o.dt
    assert same_bits(d.state(), o.state())


@pytest.mark.parametrize("bc", [(2,) * 6, (1, 1, 2, 2, 0, 1), (2, 2, 1, 1, 1, 2)])
def test_z_face_starting_a_tile_bitwise(bc):
    """nz = 65: the last 64-cell z tile holds a single cell, so the z-max
    face falls on a tile's first lane; mixed outflow, reflective and
    periodic faces."""
    shape = (12, 14, 65)
    g = Grid(*shape, G=0.0, gamma=1.4, cfl=0.35, bc=bc)
    U0 = random_smooth(shape, 1.4, seed=11)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*shape, G=0.0, gamma=1.4, cfl=0.35, bc=bc)
    d.set_state(U0)
    for _ in range(4):
        o.step()
        d.step()
        assert d.dt == o.dt
    assert same_bits(d.state(), o.state())


@pytest.mark.parametrize("nz", [66, 67, 130])
@pytest.mark.parametrize("bc", [(2,) * 6, (1, 0, 0, 1, 2, 2)])
def test_z_outflow_faces_from_step_bitwise(nz, bc):
    """Outflow on both z sides: the step kernel writes the z faces from the
    neighbouring lane's new state.  Here the last z tile is short (the z-max
    face on its lane 1 or 2), next to reflective, periodic and outflow x and
    y faces that the boundary kernel still finalises."""
    shape = (9, 11, nz)
    g = Grid(*shape, G=0.0, gamma=5 / 3, cfl=0.35, bc=bc)
    U0 = random_smooth(shape, 5 / 3, seed=nz)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*shape, G=0.0, gamma=5 / 3, cfl=0.35, bc=bc)
    d.set_state(U0)
    for _ in range(4):
        o.step()
        d.step()
        assert d.dt == o.dt
    assert same_bits(d.state(), o.state())
