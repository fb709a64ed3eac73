"""The device's sequential sums (hydro_sequential_sum through the C ABI):
the plain add chain, the peak-chunk method and its checked form all equal
an in-order loop on the CPU bit for bit, on the sequences of
tests/seqsum_model.py that stress the method (ties, walks across zero,
cancellation, wide dynamic range, subnormals, zeros, ragged lengths, column
swings); method 3 is the peak-chunk method with the LDS-staged tabulation
that the 53 moment sums of a solve take.  The checked form re-runs every
chunk from its exact start: no
re-run misses (the certificates hold), and the counters show the tables
carrying the sum (few chunks added term by term, few groups walked chunk by
chunk)."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
import seqsum_model as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [0, 1, 2, 3])
@pytest.mark.parametrize("name", list(M.hard_cases().keys()))
def test_sequential_sum_bitwise(name, method):
    print(f"sequential sum {name} method {method}", flush=True)
    x = M.hard_cases()[name]
    got = H.sequential_sum(x, method)
    want = M.plain_sum(x)
    assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64), (got, want)


@pytest.mark.parametrize("name", ["positive uniform", "mixed normal", "column swings",
                                  "halves (ties)"])
def test_peak_chunks_carry_the_sum(name):
    x = M.hard_cases()[name]
    got, st = H.sequential_sum(x, 2, stats=True)
    print(name, st)
    assert np.float64(got).view(np.int64) == np.float64(M.plain_sum(x)).view(np.int64)
    assert st["verify_failed"] == 0 and st["chain_redo"] == 0
    assert st["chunks_added"] <= 0.02 * st["chunks"]


@pytest.mark.parametrize("method", [2, 3])
def test_peak_chunks_on_many_blocks(method):
    """Four million terms of column swings (1024 chunks) against the CPU loop,
    checked (method 2) and with the LDS-staged tabulation (method 3)."""
    rng = np.random.default_rng(3)
    n = 1 << 22
    col = np.tile(np.linspace(-1.0, 1.0, 256), n // 256)
    x = col * (1.0 + 0.01 * rng.standard_normal(n)) * 1e-13 + 1e-19
    got, st = H.sequential_sum(x, method, stats=True)
    want = 0.0
    for v in np.split(x, 64):
        want = M.plain_sum(np.concatenate([[want], v]))
    print(st)
    assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64)
    assert st["chain_redo"] == 0 and st["verify_failed"] == 0
    assert st["groups_walked"] <= 4  # of 64 groups of 16 chunks


@pytest.mark.parametrize("name", ["signed zeros", "zeros between terms"])
def test_zero_chunks_pass_the_sum_through(name):
    x = M.hard_cases()[name]
    got, st = H.sequential_sum(x, 2, stats=True)
    print(name, st, flush=True)
    assert np.float64(got).view(np.int64) == np.float64(M.plain_sum(x)).view(np.int64)
    assert st["chunks_zero"] > 0 and st["verify_failed"] == 0 and st["chain_redo"] == 0
    if name == "signed zeros":
        assert st["chunks_zero"] == st["chunks"] and st["chunks_added"] == 0
