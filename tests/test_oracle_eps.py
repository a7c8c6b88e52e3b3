"""Pins the oracle's approximate search (eps > 0) by what the mathematics fixes.

The reference prunes a subtree when `lower * eps_factor > bound`, with
eps_factor = reduce(1 + eps) and the bound never below the final k-th
distance (src/query.rs:58-62, :242-244; src/metric.rs:63-71).  So every
point it never visits lies farther than r_k / (1 + eps), r_k the k-th
returned distance.  Hence, with t_i the true i-th nearest distance, the i-th
returned distance satisfies r_i <= (1 + eps) t_i for every i: either the true
i nearest were all visited (then r_i <= t_i, the returned list being the
exact top-k of the visited points), or one was not (then t_i >= its distance
> r_k / (1 + eps) >= r_i / (1 + eps)).  Checked here against brute force on
uniform and lattice data for p in {1, 2, inf}, together with: each returned
distance is the exact distance of the index it names (d = 3: (a0 + a1) + a2,
the reference's Packed d3 order), rows sorted by (distance, index) with no
index twice, and eps = 0 equal to brute force.  The checks have teeth: an
oracle run with a wrong eps factor breaks the bound, and eps >= 0.5 does
return other rows than the exact search on these inputs.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import OracleTree

K = 8
SLACK = 1.0 + 1e-12  # rounding of the reduced gates and of sqrt


def exact_dist(data, queries, p):
    a = np.abs(data[None, :, :] - queries[:, None, :])
    if p == 1.0:
        return (a[..., 0] + a[..., 1]) + a[..., 2]
    if p == 2.0:
        a = a * a
        return np.sqrt((a[..., 0] + a[..., 1]) + a[..., 2])
    return np.maximum(np.maximum(a[..., 0], a[..., 1]), a[..., 2])


def truth(dist, k):
    n = dist.shape[1]
    order = np.stack([np.lexsort((np.arange(n), row))[:k] for row in dist])
    return np.take_along_axis(dist, order, 1), order


def violations(od, oi, dist, eps):
    """Rows breaking any property of an eps-approximate k-NN answer."""
    td, _ = truth(dist, od.shape[1])
    bad = set()
    for r in range(od.shape[0]):
        d, i = od[r], oi[r]
        if len(set(i.tolist())) != len(i) or i.min() < 0 or i.max() >= dist.shape[1]:
            bad.add(r)
            continue
        if dist[r, i].tobytes() != d.tobytes():  # the exact distance of the named point
            bad.add(r)
        dd, di = d[1:] - d[:-1], i[1:] - i[:-1]
        if not np.all((dd > 0) | ((dd == 0) & (di > 0))):
            bad.add(r)
        if np.any(d > (1.0 + eps) * td[r] * SLACK):
            bad.add(r)
    return bad


def inputs(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(size=(3000, 3)), rng.uniform(size=(300, 3))
    data = rng.integers(-6, 7, size=(3000, 3)).astype(np.float64)  # lattice: ties everywhere
    return data, rng.integers(-12, 13, size=(300, 3)).astype(np.float64) / 2.0


@pytest.mark.parametrize("kind", ["uniform", "lattice"])
@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
@pytest.mark.parametrize("eps", [0.1, 0.5, 1.0])
def test_eps_answers_are_within_the_bound(kind, p, eps):
    data, queries = inputs(kind, 40 + int(10 * eps))
    dist = exact_dist(data, queries, p)
    for leafsize in (2, 16):
        od, oi = OracleTree(data, leafsize).query(queries, k=K, p=p, eps=eps)
        assert not violations(od, oi, dist, eps), (kind, p, eps, leafsize)


@pytest.mark.parametrize("kind", ["uniform", "lattice"])
@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
def test_eps_zero_is_exact(kind, p):
    data, queries = inputs(kind, 7)
    dist = exact_dist(data, queries, p)
    td, ti = truth(dist, K)
    od, oi = OracleTree(data, 4).query(queries, k=K, p=p, eps=0.0)
    np.testing.assert_array_equal(oi, ti)
    assert od.tobytes() == td.tobytes()


@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
def test_bound_rejects_a_wrong_eps_factor(p):
    # A planted mistake in the oracle's use: answers pruned with a factor of
    # 1 + 4 eps + 3 instead of 1 + eps must break the (1 + eps) bound.
    data, queries = inputs("uniform", 99)
    dist = exact_dist(data, queries, p)
    eps = 0.1
    od, oi = OracleTree(data, 2).query(queries, k=K, p=p, eps=4 * eps + 3)
    assert violations(od, oi, dist, eps), p


@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
def test_eps_path_is_exercised(p):
    # The approximation changes answers on these inputs, so the bound check
    # above is not vacuous.
    data, queries = inputs("uniform", 45)
    dist = exact_dist(data, queries, p)
    _, ti = truth(dist, K)
    od, oi = OracleTree(data, 2).query(queries, k=K, p=p, eps=1.0)
    assert np.any(oi != ti)
