"""The reference's initial-condition problems (hydro_solver_3d.c:927-2133),
made on the host by hydro_fill_initial_state, against the compiled
reference: the SHA-256 of each state in tests/golden/ic_problems.npz (from
tests/golden/make_ic_fixture.py), and the reference's own state directly
when oracle/_ref is built.  Cases: Kelvin-Helmholtz, blast wave, stellar
profile, binary (one profile; two profiles with an automatic circular orbit
and with given positions and velocities), octant explosion (thermal;
kinetic with velocity, density and spherical-harmonic perturbations and a
torus, on a box reaching below the octant)."""
import hashlib
import os
import sys

import numpy as np
import pytest

import hydrocode_astro1_amd as H
import oracle_lib as OL

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import ic_cases  # noqa: E402
from make_ic_fixture import reference_state  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "ic_problems.npz"))


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    d = tmp_path_factory.mktemp("ic")
    ic_cases.prepare(str(d))
    return str(d)


@pytest.mark.parametrize("case", ic_cases.CASES, ids=[c[0] for c in ic_cases.CASES])
def test_initial_state_equals_the_reference(case, workdir, monkeypatch):
    name, problem, shape, box, gamma, G, deck = case
    monkeypatch.chdir(workdir)
    path = None
    if deck:
        path = os.path.join(workdir, name + ".deck")
        with open(path, "w") as f:
            f.write(deck)
    U = H.initial_state(problem if problem is not None else 0, shape, box, gamma, G, path)
    assert U is not None and np.all(np.isfinite(U))
    digest = np.frombuffer(hashlib.sha256(U.tobytes()).digest(), dtype=np.uint8)
    assert np.array_equal(digest, FIX[name])
    ref = OL.ref_lib()
    if ref is not None:
        R = reference_state(ref, problem, shape, box, gamma, G, path)
        assert np.array_equal(U.view(np.int64), R.view(np.int64))


def test_problem_without_generator_leaves_the_state():
    assert H.initial_state(0, (4, 4, 4)) is None  # KELVIN_HELMHOLTZ (2D): no 3D generator


def test_profile_file_is_a_polytrope(tmp_path):
    """The profile writer's star: density falls from 1 at the centre to 0 at
    R, internal energy u = K rho^(1/n) / (gamma - 1) (n = 1.5), and the
    reference's mass integral 4 pi int rho r^2 dr matches the Lane-Emden
    mass 4 pi alpha^3 rho_c xi1^2 |theta'(xi1)|."""
    from hydrocode_astro1_amd import ics
    p = ics.write_profile_file(str(tmp_path / "p.txt"), nidx=1.5, R=1.0, points=4000)
    r, rho, u = np.loadtxt(p).T
    assert rho[0] == 1.0 and rho[-1] == 0.0 and abs(r[-1] - 1.0) < 1e-12
    assert np.all(np.diff(rho) <= 0.0)
    k = u[1:-1] / rho[1:-1] ** (1.0 / 1.5)
    assert np.allclose(k, k[0], rtol=1e-8)
    _, _, xi1, w = ics.lane_emden(1.5)
    m = np.sum(4.0 * np.pi * 0.5 * (rho[1:] + rho[:-1]) * (0.5 * (r[1:] + r[:-1])) ** 2 * np.diff(r))
    assert m == pytest.approx(4.0 * np.pi * (1.0 / xi1) ** 3 * w, rel=1e-4)
