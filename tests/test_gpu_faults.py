"""Fault paths on the device against the CPU oracle (tests/fault_cases.py):
states with NaN, infinite, negative and sub-floor cells, set through
solver3d_set_state and through a restart CSV.  The dt of every step follows
the reference's ternary-MAX fold over the speeds in cell order
(compute_dt_3d, hydro_solver_3d.c:422-450: the running maximum restarts
after the last NaN speed, a NaN last speed gives a NaN dt), and the states
are compared bit for bit where they hold numbers and NaN for NaN (NaN
payloads differ between processors and carry nothing)."""
import ctypes as C

import numpy as np
import pytest

import fault_cases as FC
import hydrocode_astro1_amd as H
import oracle_lib as OL
from oracle_lib import Grid, OracleSolver

pytestmark = pytest.mark.gpu

BC = (1, 2, 0, 0, 2, 1)  # reflective, outflow and periodic faces


@pytest.mark.parametrize("name", FC.NAMES)
def test_fault_state_steps_match_oracle(name):
    U0 = FC.case(name)
    print(f"[faults] {name}", flush=True)
    g = Grid(*FC.SHAPE, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=BC)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*FC.SHAPE, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=BC)
    d.set_state(U0)
    for step in range(3):
        o.step()
        d.step()
        assert FC.same_bits_or_nan([d.dt], [o.dt]), (name, step, d.dt, o.dt)
        assert FC.same_bits_or_nan(d.state(), o.state()), (name, step)
    d.close()


def test_nan_dt_after_many_blocks():
    """A grid with many boundary-kernel blocks and x chunks: a NaN momentum
    near the end of the cell order, a fast cell before it; the step after
    the one that spreads the NaN must take the reference's fold."""
    shape = (40, 36, 70)
    U0 = FC.smooth(shape)
    U0[37, 30, 60, 1] = float("nan")
    U0[5, 5, 5, 1] = 25.0 * U0[5, 5, 5, 0]
    g = Grid(*shape, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=(2,) * 6)
    o = OracleSolver(g)
    o.set_state(U0)
    d = H.Solver(*shape, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=(2,) * 6)
    d.set_state(U0)
    for step in range(3):
        o.step()
        d.step()
        assert FC.same_bits_or_nan([d.dt], [o.dt]), (step, d.dt, o.dt)
    assert FC.same_bits_or_nan(d.state(), o.state())
    d.close()


def test_restart_csv_with_nan_and_negative_rows(tmp_path, monkeypatch):
    """NaN and negative primitives read from a restart CSV (the reference's
    reader takes "nan" as a number) reach the state as the compiled
    reference's reader leaves them, and the steps that follow match the
    oracle."""
    R = OL.ref_lib()
    if R is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built")
    R.load_from_csv_restart.argtypes = [C.c_void_p, C.c_char_p]
    R.load_from_csv_restart.restype = C.c_int
    shape = FC.SHAPE
    g = Grid(*shape, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=BC)
    X, Y, Z = OL.cell_centres(g)
    rng = np.random.default_rng(5)
    rows = ["x,y,z,rho,u,v,w,p,gx,gy,gz"]
    for idx in np.ndindex(shape):
        rho, u, v, w, p = 1.0 + 0.1 * rng.random(), 0.1, -0.2, 0.05, 1.0
        if idx == (3, 4, 5):
            u = float("nan")
        if idx == (7, 1, 2):
            rho = -0.25
        if idx == (9, 8, 6):
            p = -2.0
        if idx == (11, 9, 7):
            v = float("nan")
        rows.append(f"{X[idx]:.6f},{Y[idx]:.6f},{Z[idx]:.6f},{rho:.6e},{u:.6e},{v:.6e},"
                    f"{w:.6e},{p:.6e},0,0,0")
    path = tmp_path / "faults.csv"
    path.write_text("\n".join(rows) + "\n")
    monkeypatch.chdir(tmp_path)
    r = OL.RefSolver(g)
    assert R.load_from_csv_restart(r.h, str(path).encode()) == 0
    d = H.Solver(*shape, G=0.0, gamma=FC.GAMMA, cfl=0.4, bc=BC)
    assert H.lib().load_from_csv_restart(d.h, str(path).encode()) == 0
    Ur = r.state()
    assert FC.same_bits_or_nan(d.state(), Ur)
    o = OracleSolver(g)
    o.set_state(Ur)
    for step in range(3):
        o.step()
        d.step()
        assert FC.same_bits_or_nan([d.dt], [o.dt]), (step, d.dt, o.dt)
        assert FC.same_bits_or_nan(d.state(), o.state()), step
    d.close()
