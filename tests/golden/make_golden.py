"""Write the SciPy-derived golden vectors used by tests/test_oracle.py.

SciPy's KDTree is the oracle the reference's own tests use
(tests/test_kdtree.py:10, :67-74 of the reference).  Inputs are regenerated in
the tests from the same numpy seeds, so only the expected outputs are stored:

  c1.npz         the plain CPU-runnable configuration: n=10,000 d=3 uniform
                 (seed 0), 1,000 uniform queries (seed 1), leafsize 32, k=8
  valid.npz      the reference's VALID cases (tests/test_kdtree.py:98-164)

Run:  python tests/golden/make_golden.py
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy
from scipy.spatial import KDTree as SciPyKDTree

HERE = Path(__file__).resolve().parent


def valid_cases():
    """The reference's VALID dict, as (name, data, queries, k, p, max_distance, leafsize)."""
    cases = [
        ("single-query", np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]), np.array([0.2, 0.0]), 2, 2.0, None, 32),
        ("batch-query", np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]), np.array([[0.2, 0.0], [3.8, 0.0]]), 2, 2.0, None, 32),
        ("max-distance-padding", np.array([[0.0, 0.0], [5.0, 0.0], [10.0, 0.0]]), np.array([0.0, 0.0]), 3, 2.0, 1.0, 32),
        ("infinite-norm", np.array([[0.0, 0.0], [2.0, 1.0], [1.0, 2.0]]), np.array([[0.1, 0.8]]), 3, np.inf, None, 1),
        ("single-query-k1", np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]), np.array([0.2, 0.0]), 1, 2.0, None, 32),
        ("batch-query-k1", np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]), np.array([[0.2, 0.0], [3.8, 0.0]]), 1, 2.0, None, 32),
        ("k-greater-than-points", np.array([[0.0], [2.0]]), np.array([[0.5]]), 4, 2.0, None, 32),
        ("promoted-dtype", np.array([[0, 0], [1, 0], [4, 0]], dtype=np.int64), np.array([[0.2, 0.0]], dtype=np.float32), 2, 2.0, None, 32),
        ("non-contiguous", np.asfortranarray([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]), np.array([[0.2, 9.0, 0.0], [3.8, 9.0, 0.0]])[:, ::2], 2, 2.0, None, 32),
        ("large-parallel-batch", np.random.default_rng(0).normal(size=(2_000, 8)), np.random.default_rng(1).normal(size=(256, 8)), 4, 2.0, None, 32),
    ]
    for dims in range(1, 18):
        for p in (1.0, 2.0, np.inf):
            name = f"row-width-{dims:02d}-p{'inf' if np.isinf(p) else int(p)}"
            cases.append(
                (
                    name,
                    np.random.default_rng(dims).normal(size=(512, dims)),
                    np.random.default_rng(1_000 + dims).normal(size=(64, dims)),
                    3,
                    p,
                    None,
                    6,
                )
            )
    return cases


def scipy_query(data, queries, k, p, max_distance, leafsize):
    tree = SciPyKDTree(np.asarray(data, dtype=np.float64), leafsize=leafsize)
    d, i = tree.query(
        np.asarray(queries, dtype=np.float64),
        k=k,
        p=p,
        distance_upper_bound=np.inf if max_distance is None else max_distance,
    )
    shape = (k,) if np.ndim(queries) == 1 else (np.shape(queries)[0], k)
    return np.asarray(d, np.float64).reshape(shape), np.asarray(i, np.int64).reshape(shape)


def c1_inputs():
    data = np.random.default_rng(0).uniform(size=(10_000, 3))
    queries = np.random.default_rng(1).uniform(size=(1_000, 3))
    return data, queries


def main():
    data, queries = c1_inputs()
    d, i = scipy_query(data, queries, 8, 2.0, None, 32)
    np.savez_compressed(HERE / "c1.npz", dist=d, idx=i, scipy_version=scipy.__version__)
    out = {}
    for name, dat, q, k, p, md, ls in valid_cases():
        d, i = scipy_query(dat, q, k, p, md, ls)
        out[f"{name}/dist"] = d
        out[f"{name}/idx"] = i
    np.savez_compressed(HERE / "valid.npz", scipy_version=scipy.__version__, **out)
    print("wrote", HERE / "c1.npz", HERE / "valid.npz")


if __name__ == "__main__":
    main()
