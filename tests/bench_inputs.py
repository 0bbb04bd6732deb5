"""Independent restatement of the throughput benchmark's input definition
(reference benchmarks/bench_ozaki2.cpp:35-56) for the tests: the splitmix64 output
function on Python integers and the Box-Muller phiX entries in numpy."""
import numpy as np

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def splitmix64(z):
    z = (z + GAMMA) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def unit53(h):
    return ((h >> 11) + 0.5) * 2.0 ** -53


def phi_matrix(rows, cols, phi, seed):
    """column-major rows x cols entries u3 exp(phi g), flattened"""
    out = np.empty(rows * cols)
    for i in range(rows * cols):
        base = i * 3 + seed
        u1, u2 = unit53(splitmix64(base)), unit53(splitmix64(base + 1))
        u3 = unit53(splitmix64(base + 2)) - 0.5
        g = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
        out[i] = u3 * np.exp(phi * g)
    return out
