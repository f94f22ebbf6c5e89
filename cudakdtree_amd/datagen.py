"""Deterministic synthetic inputs for tests and benchmarks.

Every value derives from SplitMix64 so the same seed gives the same points
on any machine and numpy version:
  uniform(n, d, seed)      -> float32 in [0,1) with 24-bit resolution
  gmm(n, d, seed, ...)     -> Gaussian mixture (Box-Muller in float64)
  unique_ints(n, d, seed)  -> a shuffled permutation of 0..n*d-1, the
                              tie-free inputs of
                              testing/testBuildersSameResult.cu:96-109
  msvc_rand_points(...)    -> the MSVC rand() sequence behind the
                              testing/issue5.cu known answer
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, count):
    """count consecutive SplitMix64 outputs for a 64-bit seed."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + _GOLDEN * np.arange(1, count + 1, dtype=np.uint64)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(n, d, seed, lo=0.0, hi=1.0):
    u = (splitmix64(seed, n * d) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))
    v = lo + (hi - lo) * u
    return v.astype(np.float32).reshape(n, d)


def gaussian(n, seed):
    u = (splitmix64(seed, 2 * n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def gmm(n, d, seed, clusters=64, sigma=0.01):
    centers = uniform(clusters, d, seed ^ 0x5151).astype(np.float64)
    which = (splitmix64(seed ^ 0xC1C1, n) % np.uint64(clusters)).astype(np.int64)
    noise = gaussian(n * d, seed ^ 0x7A7A).reshape(n, d)
    return (centers[which] + sigma * noise).astype(np.float32)


def unique_ints(n, d, seed):
    r = splitmix64(seed, n * d)
    perm = np.argsort(r, kind="stable")
    return perm.astype(np.int32).reshape(n, d)


def msvc_rand_points(count, seed, max_val=100.0):
    """testing/issue5.cu:25-40 with the MSVC rand() LCG
    (seed = seed*214013 + 2531011; out = (seed >> 16) & 0x7fff)."""
    out = np.empty(count * 3, np.float32)
    s = seed & 0xFFFFFFFF
    scale = 2.0 * max_val / 32767.0
    for i in range(count * 3):
        s = (s * 214013 + 2531011) & 0xFFFFFFFF
        out[i] = np.float32(((s >> 16) & 0x7FFF) * scale - max_val)
    return out.reshape(count, 3)
