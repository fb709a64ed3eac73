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
    if name != "tiny after huge":
        # 1e300 cancelled by -1e300: the variants' start offsets dwarf the
        # small sums that follow, so one chunk leaves its predicted path
        # unnoticed by the stitch's checks and only the verification sees it
        assert st["verify_fail"] == 0, st


def test_moment_terms_of_a_star():
    n = 64
    U = ics.explosion_n3(n)
    rho = np.maximum(U[..., 0], 1e-5)
    X = (np.arange(n) + 0.5) / n
    x, y, z = (a.ravel() for a in np.meshgrid(X, X, X, indexing="ij"))
    dm = rho.ravel() / n ** 3
    M0 = M.plain_sum(dm)
    xc, yc, zc = (M.plain_sum(dm * c) / M0 for c in (x, y, z))
    X_, Y_, Z_ = x - xc, y - yc, z - zc
    r2 = X_ * X_ + Y_ * Y_ + Z_ * Z_
    for name, t in {"M": dm, "Mx": dm * x, "Qxx": dm * (3.0 * X_ * X_ - r2),
                    "Qxz": dm * (3.0 * X_ * Z_), "Qyz": dm * (3.0 * Y_ * Z_)}.items():
        st = {}
        assert same(M.sequential_sum(t, st), M.plain_sum(t)), name
        assert st["verify_fail"] == 0 and st["fallback"] <= 0.02 * st["chunks"], (name, st)


@pytest.mark.parametrize("name", ["positive uniform", "mixed normal", "halves (ties)",
                                  "ragged length", "column swings"])
def test_a_wrong_guess_costs_time_not_bits(name):
    """Predicted peaks spoiled on purpose (each scaled by a random factor in
    [1/4, 4], a third of them negated): chunks fall back or fail the
    verification, and the result is still the plain sum."""
    rng = np.random.default_rng(11)
    x = M.hard_cases()[name]

    def noise(m):
        f = np.exp2(rng.uniform(-2.0, 2.0, m))
        return np.where(rng.random(m) < 1 / 3, -f, f)

    st = {}
    assert same(M.sequential_sum(x, st, guess_noise=noise), M.plain_sum(x)), st
    assert st["fallback"] > 0 or st["verify_fail"] > 0


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
