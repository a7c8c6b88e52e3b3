"""Parity at the full benchmark sizes, through properties that hold at any size
(the oracle's own check runs on samples of these workloads in bench.py):

* the tree: the exported indices are a permutation of 0..n-1, every tree row is
  the input row it names, bit for bit, and the `data` getter returns the input;
* every result row is sorted by (distance, index) with no index twice, and every
  returned distance is the reference's FP64 distance of that (query, point) pair,
  recomputed here in its summation order for the width (src/kernel.rs Packed d3:
  (a0 + a1) + a2; d8 full chunks: ((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7)) -- every C3
  leaf holds 32 rows, four full chunks);
* on a random sample of queries, exhaustive search over all points (exact numpy
  arithmetic in the same order) gives the same k nearest, ties by index;
* a second call, and the same queries in another order, give the same rows; on
  C3 the FP32 screen gives the same bytes as the FP64 path.

Plain GPU kernel workload: 2^20 uniform points in [0,1)^3, 2^20 queries, k = 16,
leafsize 32.  Precision workload: 2^22 8-d Gaussian-mixture points, 10^6 queries,
k = 32, leafsize 32.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def d3(q, p):
    a = (p - q) ** 2
    return np.sqrt((a[..., 0] + a[..., 1]) + a[..., 2])


def d8(q, p):
    a = (p - q) ** 2
    return np.sqrt(((a[..., 0] + a[..., 1]) + (a[..., 2] + a[..., 3])) +
                   ((a[..., 4] + a[..., 5]) + (a[..., 6] + a[..., 7])))


def check_tree(tree, data):
    tdata, idx, _, _ = tree._export()
    n = data.shape[0]
    assert np.array_equal(np.sort(idx.astype(np.int64)), np.arange(n))
    assert tdata.tobytes() == data[idx.astype(np.int64)].tobytes()
    assert tree.data.tobytes() == data.tobytes()


def check_rows(data, queries, dist, idx, dfun, chunk=65536):
    nq, k = idx.shape
    assert idx.min() >= 0 and idx.max() < data.shape[0]
    # sorted by (distance, index); no index twice in a row
    dd, di = dist[:, 1:] - dist[:, :-1], idx[:, 1:] - idx[:, :-1]
    assert np.all((dd > 0) | ((dd == 0) & (di > 0)))
    for s in range(0, nq, chunk):
        e = min(nq, s + chunk)
        ref = dfun(queries[s:e, None, :], data[idx[s:e]])
        assert ref.tobytes() == dist[s:e].tobytes(), f"rows {s}..{e}"


def check_brute(data, queries, dist, idx, dfun, rows):
    for r in rows:
        all_d = dfun(queries[r][None, :], data)
        order = np.lexsort((np.arange(data.shape[0]), all_d))[: idx.shape[1]]
        np.testing.assert_array_equal(idx[r], order, err_msg=f"query {r}")
        assert all_d[order].tobytes() == dist[r].tobytes()


def test_plain_workload_full_size():
    from kdtree import KDTree
    from kdtree.workloads import C2, c2_inputs

    data, queries = c2_inputs()
    k = C2["k"]
    print(f"n={data.shape[0]} nq={queries.shape[0]} k={k}", flush=True)
    tree = KDTree(data, leafsize=C2["leafsize"])
    check_tree(tree, data)
    dist, idx = tree.query(queries, k=k)
    check_rows(data, queries, dist, idx, d3)
    print("rows checked", flush=True)
    rng = np.random.default_rng(7)
    check_brute(data, queries, dist, idx, d3, rng.choice(queries.shape[0], 48, replace=False))
    d2, i2 = tree.query(queries, k=k)
    assert d2.tobytes() == dist.tobytes() and np.array_equal(i2, idx)
    perm = rng.permutation(queries.shape[0])[: 1 << 16]
    d3p, i3p = tree.query(queries[perm], k=k)
    assert d3p.tobytes() == dist[perm].tobytes() and np.array_equal(i3p, idx[perm])


def test_precision_workload_full_size():
    from kdtree import KDTree
    from kdtree.workloads import C3, c3_inputs

    data, queries = c3_inputs()
    k = C3["k"]
    print(f"n={data.shape[0]} nq={queries.shape[0]} k={k}", flush=True)
    tree = KDTree(data, leafsize=C3["leafsize"])
    assert data.shape[0] % C3["leafsize"] == 0 and tree.n_points == data.shape[0]
    check_tree(tree, data)
    dist, idx = tree.query(queries, k=k)
    print("queried", flush=True)
    rng = np.random.default_rng(8)
    sample = np.sort(rng.choice(queries.shape[0], 1 << 16, replace=False))
    check_rows(data, queries[sample], dist[sample], idx[sample], d8, chunk=16384)
    print("rows checked", flush=True)
    check_brute(data, queries, dist, idx, d8, rng.choice(queries.shape[0], 8, replace=False))
    print("exhaustive sample checked", flush=True)
    ds, is_ = tree.query(queries, k=k, fp32_screen=True)
    assert ds.tobytes() == dist.tobytes() and np.array_equal(is_, idx)
