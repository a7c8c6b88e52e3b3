"""Synthetic inputs for the benchmark configurations.

Values come from a SplitMix64 stream with the recipe of the reference's own
benchmark generator (benches/grid.rs:33-60): uniform = (x >> 11) * 2^-53 and a
Box-Muller normal that consumes two uniforms and keeps the cosine branch.
Draw i of a stream with seed s is SplitMix64's (i+1)-th output, so any slice
of a stream is computed directly.
"""

from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start .. start+count-1 of SplitMix64(seed)."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, start: int, count: int) -> np.ndarray:
    return (splitmix64(seed, start, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal_from(u: np.ndarray) -> np.ndarray:
    """Box-Muller over consecutive uniform pairs (u0, u1): benches/grid.rs:55-59."""
    a = 1.0 - u[0::2]
    b = u[1::2]
    return np.sqrt(-2.0 * np.log(a)) * np.cos(2.0 * np.pi * b)


def uniform_points(seed: int, n: int, d: int) -> np.ndarray:
    return uniform(seed, 0, n * d).reshape(n, d)


# Plain GPU kernel configuration: 2^20 points and 2^20 queries, d = 3,
# uniform on [0,1)^3, leafsize 32, k = 16, p = 2.
C2 = dict(n=1 << 20, nq=1 << 20, d=3, leafsize=32, k=16, data_seed=2, query_seed=1002)


def c2_inputs(n: int = C2["n"], nq: int = C2["nq"]):
    return uniform_points(C2["data_seed"], n, C2["d"]), uniform_points(C2["query_seed"], nq, C2["d"])


# Precision configuration: 2^22 points, d = 8, 64-component isotropic Gaussian
# mixture (centres uniform on [0,10)^8, sigma 0.3), 10^6 queries, k = 32.
C3 = dict(n=1 << 22, nq=1_000_000, d=8, leafsize=32, k=32, centres=64, sigma=0.3, data_seed=3,
          query_seed=1003)


def _mixture(seed: int, start: int, n: int, centres: np.ndarray, sigma: float) -> np.ndarray:
    ncent, d = centres.shape
    per = 1 + 2 * d  # component draw, then d normals of two uniforms each
    u = uniform(seed, start, n * per).reshape(n, per)
    comp = np.minimum((u[:, 0] * ncent).astype(np.int64), ncent - 1)
    z = normal_from(u[:, 1:].reshape(-1)).reshape(n, d)
    return centres[comp] + sigma * z


def c3_centres() -> np.ndarray:
    return 10.0 * uniform(C3["data_seed"], 0, C3["centres"] * C3["d"]).reshape(C3["centres"], C3["d"])


def c3_inputs(n: int = C3["n"], nq: int = C3["nq"]):
    cent = c3_centres()
    data = _mixture(C3["data_seed"], cent.size, n, cent, C3["sigma"])
    queries = _mixture(C3["query_seed"], 0, nq, cent, C3["sigma"])
    return data, queries


# Batching configuration: 256 independent trees, n log-uniform in [1e3, 1e6],
# d uniform from {2, 3, 8, 16}, float32 uniform points (promoted to float64),
# per-tree query counts log-uniform in [1, 1e5], k from {1, 8, 32}.
C4 = dict(n_trees=256, seed=4, dims=(2, 3, 8, 16), ks=(1, 8, 32))


def uniform_f32(seed: int, start: int, count: int) -> np.ndarray:
    """float32 uniforms (x >> 40) * 2^-24, promoted to float64."""
    return ((splitmix64(seed, start, count) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24)))


def c4_specs(n_trees: int = C4["n_trees"], scale: float = 1.0):
    """(n, d, n_queries, k) per tree; `scale` < 1 shrinks n and n_queries for tests."""
    u = uniform(C4["seed"], 0, 4 * n_trees).reshape(n_trees, 4)
    specs = []
    for t in range(n_trees):
        n = max(2, int(round(10 ** (3 + 3 * u[t, 0]) * scale)))
        d = C4["dims"][min(3, int(u[t, 1] * 4))]
        q = max(1, int(round(10 ** (5 * u[t, 2]) * scale)))
        k = C4["ks"][min(2, int(u[t, 3] * 3))]
        specs.append((n, d, q, k))
    return specs


def c4_tree_inputs(t: int, n: int, d: int, q: int):
    data = uniform_f32(10_000 + t, 0, n * d).reshape(n, d)
    queries = uniform_f32(20_000 + t, 0, q * d).reshape(q, d)
    return data, queries


# Serving: 32 resident trees, n log-uniform in [1e5, 8e6], d alternating
# 16 / 64, float32 standard normal points (promoted to float64); requests
# arrive as a Poisson stream, each on a uniform tree with a batch size
# log-uniform in [1, 65536] and k = 10.
C5 = dict(n_trees=32, seed=5, dims=(16, 64), k=10, max_batch=65536, hbm_budget_bytes=200e9)


def c5_specs(n_trees: int = C5["n_trees"], scale: float = 1.0):
    """(n, d) per tree; `scale` < 1 shrinks n for tests."""
    u = uniform(C5["seed"], 0, n_trees)
    return [(max(2, int(round(10 ** (5 + np.log10(80.0) * u[t]) * scale))), C5["dims"][t % 2])
            for t in range(n_trees)]


def c5_queries(request: int, m: int, d: int) -> np.ndarray:
    """Standard normal float32 query rows of one request, promoted to float64."""
    u = uniform(1_000_000 + request, 0, 2 * m * d)
    return normal_from(u)[: m * d].astype(np.float32).astype(np.float64).reshape(m, d)
