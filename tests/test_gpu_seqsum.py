"""The device's sequential sums (hydro_sequential_sum through the C ABI):
the plain add chain and the block-composite walk both equal an in-order
loop on the CPU bit for bit, on the sequences of tests/seqsum_model.py that
stress the composite method (ties, walks across zero, cancellation, wide
dynamic range, subnormals, zeros, ragged lengths)."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
import seqsum_model as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name", list(M.hard_cases().keys()))
def test_sequential_sum_bitwise(name, method):
    print(f"sequential sum {name} method {method}", flush=True)
    x = M.hard_cases()[name]
    got = H.sequential_sum(x, method)
    want = M.plain_sum(x)
    assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64), (got, want)
