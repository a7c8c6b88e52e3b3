"""Tile walk for rows of 1-4, 8 or 16 coordinates (p = 2, eps = 0, k <= 32).

One tree walk per wavefront of 64 queries, an FP32 screen of the points it
reaches with a proven margin, the FP64 recheck and the visit certificate; the
traversal answers the queries the walk flags.  Every result must equal the
oracle's (the reference's descent restated) bit for bit.  The screen-only
diagnostic flag skips that traversal, so the rows it writes show what the walk
itself answered (and those must be exact too).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from kdtree import KDTree
from oracle import OracleTree

pytestmark = pytest.mark.gpu

FP32_SCREEN = 0x1
SCREEN_ONLY = 0x4
TILE_WALK = 0x400
NO_TILE_WALK = 0x800


def query_flags(tree: KDTree, q: np.ndarray, k: int, md, flags: int):
    out_d = np.full((q.shape[0], k), np.nan)
    out_i = np.full((q.shape[0], k), -7, dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(q.ctypes.data), q.shape[0], q.shape[1], k, 2.0,
                             np.inf if md is None else md, 0.0, flags,
                             ctypes.c_void_p(out_d.ctypes.data), ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    return out_d, out_i


def mixture(seed: int, n: int, d: int, sigma: float, ncent: int = 16) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 10, size=(ncent, d))
    return centres[rng.integers(0, ncent, size=n)] + sigma * rng.standard_normal((n, d))


def lattice(seed: int, n: int, d: int, side: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, side, size=(n, d)).astype(np.float64)


def u(seed, n, d):
    return np.random.default_rng(seed).uniform(size=(n, d))


# name: (data, queries, k, leafsize, max_distance, (min, max) fraction the walk answers itself)
CASES = {
    "uniform-20000x3-k16": (lambda: u(1, 20_000, 3), lambda: u(2, 3000, 3), 16, 32, None, (0.99, 1.0)),
    "uniform-5000x1-k8": (lambda: u(3, 5000, 1), lambda: u(4, 777, 1), 8, 32, None, (0.99, 1.0)),
    "uniform-9000x2-k32-ls8": (lambda: u(5, 9000, 2), lambda: u(6, 1000, 2), 32, 8, None, (0.99, 1.0)),
    "uniform-7000x4-k5-ls7": (lambda: u(7, 7000, 4), lambda: u(8, 515, 4), 5, 7, None, (0.99, 1.0)),
    "mixture-30000x8-k32": (lambda: mixture(9, 30_000, 8, 0.3), lambda: mixture(10, 2000, 8, 0.3), 32, 32, None,
                            (0.97, 1.0)),
    "normal-12000x16-k10": (lambda: np.random.default_rng(11).normal(size=(12_000, 16)),
                            lambda: np.random.default_rng(12).normal(size=(700, 16)), 10, 32, None, (0.97, 1.0)),
    "normal-6000x8-k1-ls64": (lambda: np.random.default_rng(13).normal(size=(6000, 8)),
                              lambda: np.random.default_rng(14).normal(size=(333, 8)), 1, 64, None, (0.99, 1.0)),
    # exact ties everywhere: the recheck's (distance, index) order decides
    "lattice-8000x3-k16": (lambda: lattice(15, 8000, 3, 12), lambda: lattice(16, 900, 3, 12), 16, 32, None,
                           (0.5, 1.0)),
    "lattice-3000x8-k8-ls4": (lambda: lattice(17, 3000, 8, 3), lambda: lattice(18, 300, 8, 3), 8, 4, None,
                              (0.0, 1.0)),
    # 100 copies of every row: more than 64 candidates, so the traversal answers
    "duplicates-4000x3-k10": (lambda: np.repeat(u(19, 40, 3), 100, axis=0), lambda: u(20, 100, 3), 10, 32, None,
                              (0.0, 0.0)),
    "tiny-5x3-k10": (lambda: u(21, 5, 3), lambda: u(22, 70, 3), 10, 32, None, (0.99, 1.0)),
    "single-point-1x2-k1": (lambda: u(23, 1, 2), lambda: u(24, 3, 2), 1, 32, None, (0.99, 1.0)),
    "cutoff-8000x3-k16": (lambda: u(25, 8000, 3), lambda: u(26, 400, 3), 16, 32, 0.05, (0.99, 1.0)),
    "cutoff-zero-2000x2-k4": (lambda: lattice(27, 2000, 2, 30), lambda: lattice(28, 200, 2, 30), 4, 16, 0.0,
                              (0.99, 1.0)),
    "far-queries-5000x8-k4": (lambda: u(29, 5000, 8), lambda: 50.0 + u(30, 100, 8), 4, 32, None, (0.99, 1.0)),
    "offset-data-6000x3-k8": (lambda: 1e6 + u(31, 6000, 3), lambda: 1e6 + u(32, 300, 3), 8, 32, None, (0.5, 1.0)),
    "huge-coords-2000x4-k3": (lambda: 1e17 * u(33, 2000, 4), lambda: 1e17 * u(34, 200, 4), 3, 32, None,
                              (0.5, 1.0)),
    # rows beyond the FP32 range: the walk does not apply, the traversal's own
    # screen (which switches itself off there) answers every row
    "beyond-fp32-1000x3-k3": (lambda: 1e19 * u(35, 1000, 3), lambda: 1e19 * u(36, 100, 3), 3, 32, None,
                              (0.99, 1.0)),
}


@pytest.mark.parametrize("case", CASES.values(), ids=CASES.keys())
def test_walk_matches_oracle(case):
    make_data, make_q, k, leafsize, md, (lo, hi) = case
    data = np.ascontiguousarray(make_data())
    q = np.ascontiguousarray(make_q())
    print(f"walk case n={data.shape[0]} d={data.shape[1]} nq={q.shape[0]} k={k} leafsize={leafsize}", flush=True)
    od, oi = OracleTree(data, leafsize).query(q, k=k, max_distance=md)
    tree = KDTree(data, leafsize=leafsize)
    wd, wi = query_flags(tree, q, k, md, FP32_SCREEN | TILE_WALK)
    np.testing.assert_array_equal(wi, oi)
    assert wd.tobytes() == od.tobytes()
    # the traversal's own FP32 screen and the plain FP64 path agree too
    td, ti = query_flags(tree, q, k, md, FP32_SCREEN | NO_TILE_WALK)
    np.testing.assert_array_equal(ti, oi)
    assert td.tobytes() == od.tobytes()
    # rows the walk answered by itself are exact as well
    sd, si = query_flags(tree, q, k, md, FP32_SCREEN | TILE_WALK | SCREEN_ONLY)
    done = si[:, 0] != -7
    np.testing.assert_array_equal(si[done], oi[done])
    assert sd[done].tobytes() == od[done].tobytes()
    print(f"walk answered {done.mean():.3f} of the queries", flush=True)
    assert lo <= done.mean() <= hi


@pytest.mark.parametrize("k", [1, 8, 16, 24, 32])
def test_walk_leaf_order_batch(k):
    """A batch large enough for the leaf-order sort (>= 8192 queries), taken by
    the walk without the explicit flag; compared with the plain FP64 traversal
    byte for byte and with the oracle on a sample."""
    data = mixture(40 + k, 60_000, 8, 0.3, ncent=32)
    q = mixture(80 + k, 12_000, 8, 0.3, ncent=32)
    print(f"walk leaf-order batch k={k}", flush=True)
    tree = KDTree(data, leafsize=32)
    wd, wi = query_flags(tree, q, k, None, FP32_SCREEN)
    pd, pi = query_flags(tree, q, k, None, 0)
    np.testing.assert_array_equal(wi, pi)
    assert wd.tobytes() == pd.tobytes()
    sample = np.random.default_rng(k).choice(q.shape[0], 300, replace=False)
    od, oi = OracleTree(data, 32).query(q[sample], k=k)
    np.testing.assert_array_equal(wi[sample], oi)
    assert wd[sample].tobytes() == od.tobytes()
    sd, si = query_flags(tree, q, k, None, FP32_SCREEN | SCREEN_ONLY)
    done = si[:, 0] != -7
    print(f"walk answered {done.mean():.4f} of the queries", flush=True)
    assert done.mean() > 0.95


@pytest.mark.parametrize("d", [1, 2, 3, 4, 8, 16])
def test_walk_every_width_uniform(d):
    data = u(100 + d, 20_000, d)
    q = u(200 + d, 9000, d)
    print(f"walk width d={d}", flush=True)
    tree = KDTree(data, leafsize=16)
    for k in (1, 7, 32):
        wd, wi = query_flags(tree, q, k, None, FP32_SCREEN | TILE_WALK)
        pd, pi = query_flags(tree, q, k, None, 0)
        np.testing.assert_array_equal(wi, pi)
        assert wd.tobytes() == pd.tobytes()
    od, oi = OracleTree(data, 16).query(q[:200], k=32)
    np.testing.assert_array_equal(wi[:200], oi)
    assert wd[:200].tobytes() == od.tobytes()
