"""Seeded synthetic inputs for the benchmark: complex tensors whose real and imaginary parts
are N(0, 1), from SplitMix64 integers (53-bit uniforms) through Box--Muller, laid out
column-major over the shape (the i-th draw is the i-th element of the buffer), so a
benchmark input is reproducible from its seed alone."""
import numpy as np

_MASK = (1 << 64) - 1


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) for a seed."""
    state = np.uint64(seed & _MASK)
    with np.errstate(over="ignore"):
        z = state + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def complex_gaussian(seed, shape):
    """Column-major complex128 array of the given shape, N(0, 1) real and imaginary parts."""
    n = int(np.prod(shape))
    u = splitmix64(seed, 2 * n)
    u1 = ((u[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (u[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    z = r * np.cos(2.0 * np.pi * u2) + 1j * (r * np.sin(2.0 * np.pi * u2))
    return np.reshape(z, shape, order="F")
