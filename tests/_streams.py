"""The library's seeded uniform stream (psdp_device.h, uniform01): element i
of stream `seed` is ((splitmix64(seed ^ C (i + 1)) >> 11) + 1) 2^-53, in
(0, 1].  Lanczos start vectors use elements 0..n-1; a LOBPCG start block
uses elements 0..n*m-1 in column-major order."""
import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def uniform_stream(seed, count):
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) ^ (np.uint64(0x632BE59BD9B4E019) * i)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53


def start_vector(seed, n):
    return uniform_stream(seed, n)


def start_block(seed, n, m):
    return uniform_stream(seed, n * m).reshape(n, m, order="F")
