"""The numpy statement of the peak-chunk sequential sum
(tests/seqsum_model.py) equals a plain in-order loop bit for bit on
sequences built to break it, and on the moment terms of a star-like
density, where nearly every chunk must go through the tabulated stitch (a
method that fell back to adding term by term everywhere would be correct
but pointless)."""
import numpy as np
import pytest

import seqsum_model as M
from hydrocode_astro1_amd import ics


def same(a, b):
    return np.float64(a).view(np.int64) == np.float64(b).view(np.int64)


@pytest.mark.parametrize("name", list(M.hard_cases().keys()))
def test_hard_cases_bitwise(name):
    x = M.hard_cases()[name]
    st = {}
    got = M.sequential_sum(x, st)
    assert same(got, M.plain_sum(x)), (got, M.plain_sum(x), st)
    # the certificates alone make the stitch exact: every chunk re-run from
    # its exact start lands on the next start
    assert st["verify_fail"] == 0, st


def _star_terms(floor=None):
    n = 64
    U = ics.explosion_n3(n)
    rho = U[..., 0] if floor is None else np.maximum(U[..., 0], floor)
    X = (np.arange(n) + 0.5) / n
    x, y, z = (a.ravel() for a in np.meshgrid(X, X, X, indexing="ij"))
    dm = rho.ravel() / n ** 3
    M0 = M.plain_sum(dm)
    xc, yc, zc = (M.plain_sum(dm * c) / M0 for c in (x, y, z))
    X_, Y_, Z_ = x - xc, y - yc, z - zc
    r2 = X_ * X_ + Y_ * Y_ + Z_ * Z_
    return {"M": dm, "Mx": dm * x, "Qxx": dm * (3.0 * X_ * X_ - r2),
            "Qxz": dm * (3.0 * X_ * Z_), "Qyz": dm * (3.0 * Y_ * Z_)}


def test_moment_terms_of_a_star():
    """The moment terms of the explosion's star: the tables carry nearly
    every chunk (none is added term by term)."""
    for name, t in _star_terms().items():
        st = {}
        assert same(M.sequential_sum(t, st), M.plain_sum(t)), name
        assert st["verify_fail"] == 0 and st["fallback"] <= 0.02 * st["chunks"], (name, st)


def test_moment_terms_over_a_uniform_floor():
    """Density clamped to a uniform floor: the sum climbs through the floor
    cells in equal steps, so a running sum can land within a few ulps of a
    binade's edge, where no start other than the tabulated one is certified.
    Those chunks are added term by term; the result stays exact."""
    for name, t in _star_terms(1e-5).items():
        st = {}
        assert same(M.sequential_sum(t, st), M.plain_sum(t)), name
        assert st["verify_fail"] == 0, (name, st)


@pytest.mark.parametrize("name", ["positive uniform", "mixed normal", "halves (ties)",
                                  "ragged length", "column swings"])
def test_a_wrong_guess_costs_time_not_bits(name):
    """Predicted peaks spoiled on purpose (each scaled by a random factor in
    [1/4, 4], a third of them negated): chunks fall back to adding term by
    term, none is carried wrongly, and the result is still the plain sum."""
    rng = np.random.default_rng(11)
    x = M.hard_cases()[name]

    def noise(m):
        f = np.exp2(rng.uniform(-2.0, 2.0, m))
        return np.where(rng.random(m) < 1 / 3, -f, f)

    st = {}
    assert same(M.sequential_sum(x, st, guess_noise=noise), M.plain_sum(x)), st
    assert st["fallback"] > 0 and st["verify_fail"] == 0


@pytest.mark.parametrize("name", ["signed zeros", "zeros between terms"])
def test_zero_chunks_pass_the_sum_through(name):
    """Chunks whose terms are all +0 or -0 leave the sum as it is (the
    moment terms that vanish by symmetry are such sequences): none of them is
    added term by term."""
    x = M.hard_cases()[name]
    st = {}
    assert same(M.sequential_sum(x, st), M.plain_sum(x))
    assert st["zero"] > 0 and st["verify_fail"] == 0
    if name == "signed zeros":
        assert st["zero"] == st["chunks"] and st["fallback"] == 0, st


@pytest.mark.parametrize("name", ["positive uniform", "mixed normal", "halves (ties)",
                                  "big halves", "powers of two", "column swings",
                                  "zeros between terms", "subnormals", "tiny after huge"])
def test_group_maps_carry_the_sum(name):
    """Groups of 4 chunks (13 groups on 200 000 terms): the walk from group
    to group through the maps gives the plain sum bit for bit, and on the
    well-behaved sequences nearly every group goes through its map."""
    x = M.hard_cases()[name]
    st = {}
    assert same(M.sequential_sum(x, st, group=4), M.plain_sum(x)), st
    if name not in ("subnormals", "tiny after huge"):
        assert st["walked"] <= 1, st


@pytest.mark.parametrize("group", [2, 8])
def test_group_maps_on_moment_terms(group):
    for name, t in _star_terms().items():
        st = {}
        assert same(M.sequential_sum(t, st, group=group), M.plain_sum(t)), name
        assert st["verify_fail"] == 0 and st["walked"] <= st["groups"] // 2, (name, st)
