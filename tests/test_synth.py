"""The benchmark's input generator (mpsdyn_amd.synth) draws the same tensors as the oracle's,
bit for bit, so benchmark inputs are the parity tests' inputs for the same seed."""
import numpy as np
import pytest

from mpsdyn_amd import synth
from oracle import mpsd_oracle as orc


@pytest.mark.parametrize("seed,shape", [(1, (3,)), (101, (5, 3, 7)), (2**63 + 5, (2, 4, 2, 3)), (104, (2 * 64 * 64,))])
def test_same_draws_as_oracle(seed, shape):
    a, b = synth.complex_gaussian(seed, shape), orc.complex_gaussian(seed, shape)
    assert a.shape == b.shape and a.dtype == np.complex128
    assert np.array_equal(a, b)


def test_moments():
    z = synth.complex_gaussian(7, (200000,))
    assert abs(z.real.mean()) < 0.01 and abs(z.imag.mean()) < 0.01
    assert abs(z.real.var() - 1.0) < 0.02 and abs(z.imag.var() - 1.0) < 0.02
