"""Seeded synthetic inputs shared by the GPU path, the tests and the bench.

Generator: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) -> 53-bit uniforms in [0, 1) -> Box-Muller normals.
A ``Stream`` is consumed in order: uniform k of a stream seeded with s is
``splitmix64(s + (k+1) * 0x9E3779B97F4A7C15) >> 11`` times 2**-53.  Box-Muller
takes uniforms in pairs (u1, u2) and returns sqrt(-2 log(1-u1)) cos(2 pi u2)
followed by sqrt(-2 log(1-u1)) sin(2 pi u2).

Point matrices are returned as D x N (a column per point).  Normals for a point
set are drawn point-major: the first D normals are point 0.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(counter_start, seed, n):
    k = np.arange(counter_start + 1, counter_start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    def __init__(self, seed=0):
        self.seed = np.uint64(seed)
        self.count = 0

    def uniform(self, n):
        z = splitmix64(self.count, self.seed, n)
        self.count += n
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        r = np.sqrt(-2.0 * np.log(1.0 - u[0::2]))
        th = 2.0 * np.pi * u[1::2]
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(th)
        out[1::2] = r * np.sin(th)
        return out[:n]

    def points(self, dim, n):
        """D x N standard normal points (point-major draw)."""
        return np.ascontiguousarray(self.normal(dim * n).reshape(n, dim).T)


def banana_points(n, dim=8, seed=0):
    """x1 ~ N(0,1), x_k = 0.5 x_{k-1}^2 + tanh(x_{k-2}) + 0.3 eps_k with x_0 := 0."""
    eps = Stream(seed).normal(dim * n).reshape(n, dim)
    x = np.zeros((n, dim))
    x[:, 0] = eps[:, 0]
    for k in range(1, dim):
        prev2 = x[:, k - 2] if k >= 2 else 0.0
        x[:, k] = 0.5 * x[:, k - 1] ** 2 + np.tanh(prev2) + 0.3 * eps[:, k]
    return np.ascontiguousarray(x.T)


def student_t_points(dim, n, nu=5, seed=0):
    """Independent Student-t(nu) coordinates: z / sqrt(chi2_nu / nu)."""
    s = Stream(seed)
    z = s.normal(dim * n).reshape(n, dim)
    chi = (s.normal(dim * n * nu).reshape(n, dim, nu) ** 2).sum(axis=2)
    return np.ascontiguousarray((z / np.sqrt(chi / nu)).T)
