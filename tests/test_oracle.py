"""Pins the oracle (oracle/kdtree_oracle.cpp) before it is trusted as the checker.

Each check is chosen so a plausible restatement mistake fails it:
  * SciPy golden vectors (the reference's own oracle, tests/test_kdtree.py:
    67-89): indices exactly, distances at 1e-12 -- a wrong metric, a dropped
    term, a wrong prune all change indices.
  * Closed forms (one non-zero axis, so every summation order agrees).
  * Closed forms that fix each row width's summation order: terms 1 and
    2^-53, where every plausible wrong pairing changes the last bit.
  * Brute force on integer lattices, where every distance is exact in FP64 and
    ties are everywhere: pins the (distance, index) tie rule, padding and
    max_distance independently of summation order.
  * Tree invariants on small builds, and a second, independent formulation of
    the build (two-pointer Hoare pairing plus a stable partition for the
    pivot-equal fix-up) that must give the identical tree.
  * A hypothesis property test against live SciPy, as the reference's
    test_compatibility does (tests/test_kdtree.py:33-95).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st
from scipy.spatial import KDTree as SciPyKDTree

from oracle import OracleError, OracleTree, count_nodes, point_dist

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402


# ------------------------------------------------------------ SciPy vectors
def test_c1_matches_scipy_golden():
    data, queries = make_golden.c1_inputs()
    g = np.load(GOLDEN / "c1.npz")
    d, i = OracleTree(data, 32).query(queries, k=8)
    np.testing.assert_array_equal(i, g["idx"])
    np.testing.assert_allclose(d, g["dist"], rtol=0, atol=1e-12)


VALID = make_golden.valid_cases()


@pytest.mark.parametrize("case", VALID, ids=[c[0] for c in VALID])
def test_valid_cases_match_scipy_golden(case):
    name, data, queries, k, p, md, ls = case
    g = np.load(GOLDEN / "valid.npz")
    d, i = OracleTree(np.asarray(data, np.float64), ls).query(np.asarray(queries, np.float64), k=k, p=p,
                                                              max_distance=md)
    np.testing.assert_array_equal(i, g[f"{name}/idx"])
    np.testing.assert_allclose(d, g[f"{name}/dist"], rtol=0, atol=1e-12)


# ------------------------------------------------------------- closed forms
def test_closed_forms_of_reference_literals():
    three = np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]])
    d, i = OracleTree(three).query(np.array([0.2, 0.0]), k=2)
    np.testing.assert_array_equal(i, [0, 1])
    assert d.tolist() == [abs(0.2 - 0.0), np.sqrt((1.0 - 0.2) ** 2)]
    d, i = OracleTree(three).query(np.array([[0.2, 0.0], [3.8, 0.0]]), k=2)
    np.testing.assert_array_equal(i, [[0, 1], [2, 1]])
    d, i = OracleTree(np.array([[0.0, 0.0], [5.0, 0.0], [10.0, 0.0]])).query(np.array([0.0, 0.0]), k=3,
                                                                              max_distance=1.0)
    assert d.tolist() == [0.0, np.inf, np.inf] and i.tolist() == [0, 3, 3]
    d, i = OracleTree(np.array([[0.0], [2.0]])).query(np.array([[0.5]]), k=4)
    assert d.tolist() == [[0.5, 1.5, np.inf, np.inf]] and i.tolist() == [[0, 1, 2, 2]]
    d, i = OracleTree(np.array([[0.0, 0.0], [2.0, 1.0], [1.0, 2.0]]), 1).query(np.array([[0.1, 0.8]]), k=3,
                                                                                  p=np.inf)
    assert i.tolist() == [[0, 2, 1]]
    assert d.tolist() == [[0.8, max(abs(1.0 - 0.1), abs(2.0 - 0.8)), max(abs(2.0 - 0.1), abs(1.0 - 0.8))]]


def test_sqrt_restore_is_ieee():
    pts = np.random.default_rng(5).normal(size=(300, 1))
    q = np.array([0.123456789])
    d, i = OracleTree(pts, 4).query(q, k=300)
    np.testing.assert_array_equal(d, np.sqrt((pts[i, 0] - q[0]) ** 2))


# -------------------------------------------------- brute force on lattices
def brute_force(data, queries, k, p, max_distance):
    diff = np.abs(data[None, :, :] - queries[:, None, :])
    if p == 1.0:
        dist = diff.sum(-1)
    elif p == 2.0:
        dist = np.sqrt((diff ** 2).sum(-1))
    else:
        dist = diff.max(-1)
    n = data.shape[0]
    out_d = np.full((queries.shape[0], k), np.inf)
    out_i = np.full((queries.shape[0], k), n, np.int64)
    cut = np.inf if max_distance is None else max_distance
    for r in range(queries.shape[0]):
        order = np.lexsort((np.arange(n), dist[r]))
        order = [j for j in order if dist[r, j] <= cut][:k]
        out_d[r, : len(order)] = dist[r, order]
        out_i[r, : len(order)] = order
    return out_d, out_i


@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
@pytest.mark.parametrize("leafsize", [1, 4, 32])
@pytest.mark.parametrize("max_distance", [None, 2.0])
def test_lattice_ties_against_brute_force(p, leafsize, max_distance):
    rng = np.random.default_rng(int(leafsize + 10 * (p if np.isfinite(p) else 9)))
    data = rng.integers(-4, 5, size=(700, 3)).astype(np.float64)  # many duplicates
    queries = rng.integers(-8, 9, size=(60, 3)).astype(np.float64) / 2.0
    k = 11
    od, oi = OracleTree(data, leafsize).query(queries, k=k, p=p, max_distance=max_distance)
    bd, bi = brute_force(data, queries, k, p, max_distance)
    np.testing.assert_array_equal(oi, bi)
    np.testing.assert_array_equal(od, bd)


# ------------------------------------------------------------ tree invariants
def total_key(x: np.ndarray) -> np.ndarray:
    b = x.view(np.int64)
    return b ^ ((b >> 63).view(np.uint64) >> np.uint64(1)).view(np.int64)


def count_nodes_py(n, ls):
    return 1 if n <= ls else 1 + count_nodes_py(n // 2, ls) + count_nodes_py(n - n // 2, ls)


@pytest.mark.parametrize("n,ls", [(1, 1), (2, 1), (10, 3), (1000, 32), (12345, 7), (65, 64), (100000, 13)])
def test_count_nodes_closed_form(n, ls):
    assert count_nodes(n, ls) == count_nodes_py(n, ls)


@pytest.mark.parametrize(
    "data,ls",
    [
        (np.random.default_rng(1).normal(size=(777, 3)), 5),
        (np.random.default_rng(2).integers(0, 4, size=(500, 4)).astype(float), 3),
        (np.random.default_rng(3).uniform(size=(300, 1)), 1),
        (np.concatenate([np.zeros((40, 2)), -np.zeros((40, 2))]), 3),
    ],
)
def test_build_invariants(data, ls):
    t = OracleTree(data, ls)
    rows, idx, nodes, boxes = t.export()
    n, d = data.shape
    assert sorted(idx.tolist()) == list(range(n))
    np.testing.assert_array_equal(rows, data[idx])
    assert len(nodes) == count_nodes(n, ls)

    def check(node_id, start, length):
        nd = nodes[node_id]
        block = rows[start:start + length]
        np.testing.assert_array_equal(boxes[node_id, 0], block.min(0))
        np.testing.assert_array_equal(boxes[node_id, 1], block.max(0))
        if length <= ls:
            assert nd["is_leaf"] and (nd["a"], nd["b"]) == (start, start + length)
            return
        assert not nd["is_leaf"]
        ext = boxes[node_id, 1] - boxes[node_id, 0]
        assert nd["split_dim"] == int(np.argmax(ext))  # first index of the maximum
        mid = length // 2
        keys = block[:, nd["split_dim"]]
        # split value = the mid-th order statistic under f64::total_cmp
        assert total_key(np.array([nd["split_value"]]))[0] == np.sort(total_key(keys.copy()))[mid]
        assert (keys[:mid] <= nd["split_value"]).all() and (keys[mid:] >= nd["split_value"]).all()
        assert nd["a"] == node_id + 1 and nd["b"] == node_id + 1 + count_nodes(mid, ls)
        shrink = lambda c: all(  # noqa: E731
            (boxes[c, 1, j] - boxes[c, 0, j]) < 0.6 * ext[j] for j in range(d) if j != nd["split_dim"]
        )
        want = d > 1 and length >= 64 and shrink(nd["a"]) and shrink(nd["b"])
        assert bool(nd["order_by_box"]) == want
        check(nd["a"], start, mid)
        check(nd["b"], start + mid, length - mid)

    check(0, 0, n)


def build_independent(data: np.ndarray, ls: int):
    """A second formulation of build::build (src/build.rs:19-314): the Hoare
    two-pointer pass the reference's sweeps are documented to equal
    (src/build.rs:257-259, :294-300), and the pivot-equal fix-up written as a
    stable partition of the window it touches."""
    rows = data.copy()
    idx = np.arange(data.shape[0], dtype=np.uint32)
    n_nodes = count_nodes(data.shape[0], ls)
    nodes = {}

    def rec(nid, start, length):
        block = rows[start:start + length]
        lo, hi = block.min(0), block.max(0)
        if length <= ls:
            nodes[nid] = ("leaf", start, start + length)
            return lo, hi
        dim = int(np.argmax(hi - lo))
        mid = length // 2
        keys = rows[start:start + length, dim].copy()
        pivot = keys[np.argsort(total_key(keys.copy()), kind="stable")[mid]]
        lt = int((keys < pivot).sum())
        i, j = 0, length - 1
        while True:
            while i < lt and rows[start + i, dim] < pivot:
                i += 1
            while j >= lt and rows[start + j, dim] >= pivot:
                j -= 1
            if i >= lt or j < lt:
                break
            a, b = start + i, start + j
            rows[[a, b]] = rows[[b, a]]
            idx[[a, b]] = idx[[b, a]]
            i += 1
            j -= 1
        need = mid - lt
        if need:
            # The fix-up as a map: the t-th swap exchanges place = lt+t-1 with
            # the t-th pivot-equal position s_t, so a row that is not
            # pivot-equal and sits in [lt, mid) follows P -> g(P) -> ... to
            # the first position >= mid; pivot-equal rows go to lt+t-1.
            s_t = lt + np.nonzero(rows[start + lt:start + length, dim] == pivot)[0][:need]
            g = {lt + t: int(s_t[t]) for t in range(need)}
            dest = {}
            for t, s in enumerate(s_t):
                dest[int(s)] = lt + t
            for pos in range(lt, mid):
                if pos in dest:  # pivot-equal row already in [lt, mid)
                    continue
                x = pos
                while x < mid:
                    x = g[x]
                dest[pos] = x
            src = np.array(sorted(dest), dtype=np.int64)
            dst = np.array([dest[v] for v in src], dtype=np.int64)
            rows[start + dst] = rows[start + src].copy()
            idx[start + dst] = idx[start + src].copy()
        left, right = nid + 1, nid + 1 + count_nodes(mid, ls)
        llo, lhi = rec(left, start, mid)
        rlo, rhi = rec(right, start + mid, length - mid)
        ext = hi - lo
        shr = lambda clo, chi: all(  # noqa: E731
            (chi[k] - clo[k]) < 0.6 * ext[k] for k in range(len(ext)) if k != dim
        )
        obb = len(ext) > 1 and length >= 64 and shr(llo, lhi) and shr(rlo, rhi)
        nodes[nid] = ("inner", left, right, dim, obb, pivot)
        return lo, hi

    rec(0, 0, data.shape[0])
    assert len(nodes) == n_nodes
    return rows, idx, nodes


@pytest.mark.parametrize(
    "data,ls",
    [
        # the reference's parallel_build_matches_serial_build data, src/build.rs:347-350
        ((np.arange(2000 * 5, dtype=np.uint64) * np.uint64(2_654_435_761) % np.uint64(10_007))
         .astype(float).reshape(2000, 5), 8),
        (np.random.default_rng(9).integers(0, 5, size=(600, 2)).astype(float), 2),
        (np.random.default_rng(10).normal(size=(900, 4)), 6),
    ],
)
def test_build_matches_independent_formulation(data, ls):
    rows, idx, nodes = build_independent(data, ls)
    orows, oidx, onodes, _ = OracleTree(data, ls).export()
    np.testing.assert_array_equal(oidx, idx)
    np.testing.assert_array_equal(orows, rows)
    for nid, nd in nodes.items():
        o = onodes[nid]
        if nd[0] == "leaf":
            assert o["is_leaf"] and (o["a"], o["b"]) == nd[1:]
        else:
            assert not o["is_leaf"]
            assert (o["a"], o["b"], o["split_dim"], bool(o["order_by_box"])) == nd[1:5]
            assert o["split_value"] == nd[5]


# --------------------------------------------------------------- errors
def test_oracle_error_codes():
    with pytest.raises(OracleError):
        OracleTree(np.zeros((0, 2)))
    t = OracleTree(np.zeros((3, 2)))
    for kw, code in [({"k": 0}, 3), ({"eps": -1.0}, 5), ({"max_distance": -1.0}, 6), ({"p": 0.5}, 4)]:
        with pytest.raises(OracleError) as e:
            t.query(np.zeros(2), **kw)
        assert e.value.code == code
    with pytest.raises(OracleError) as e:
        t.query(np.array([np.nan, 0.0]), eps=-1.0)
    assert e.value.code == 7  # finiteness is checked before eps (src/query.rs:46-51)


# ------------------------------------------------- per-width recipe vectors
def test_point_dist_width_recipes_are_plausible():
    # Every recipe is some order of the same terms, so all agree with the
    # plain sum to rounding; the exact order is the restatement's (src/kernel.rs).
    rng = np.random.default_rng(12)
    for d in range(1, 40):
        q, pt = rng.normal(size=d), rng.normal(size=d)
        want = ((q - pt) ** 2).sum()
        for off, length in [(0, 32), (31, 33)]:
            assert abs(point_dist(q, pt, 2.0, off, length) - want) <= 1e-13 * max(1.0, want)


from summation_cases import ONE, ONE_UP, ORDER_CASES, terms  # noqa: E402


@pytest.mark.parametrize("case", ORDER_CASES, ids=[c[0] for c in ORDER_CASES])
def test_point_dist_summation_order_closed_forms(case):
    _, d, v, off, length, want, _mistake = case
    got = point_dist(np.zeros(d), v, 1.0, off, length)
    assert got == want, f"{got!r} != {want!r}"
    # L2 with squares of powers of two: terms 1 and four of 2^-54 (two pairs
    # make 2^-52) where the L1 case relies on a pair of T.
    if d == 8 and off == 0:
        sq = terms(8, a0=1.0, a4=2.0**-27, a5=2.0**-27, a6=2.0**-27, a7=2.0**-27)
        assert point_dist(np.zeros(8), sq, 2.0, 0, 8) == ONE_UP
        assert point_dist(np.zeros(8), sq, 2.0, 8, 9) == ONE


# ---------------------------------------------------------- property test
@st.composite
def problems(draw):
    dims = draw(st.integers(1, 20))
    n_points = draw(st.integers(1, 3000))
    n_queries = draw(st.integers(1, 32))
    rng = np.random.default_rng(draw(st.integers(0, 2**32 - 1)))
    queries = rng.normal(size=(n_queries, dims))
    if draw(st.booleans()):
        queries = queries[0]
    return (
        rng.normal(size=(n_points, dims)),
        queries,
        draw(st.integers(1, 16)),
        draw(st.sampled_from((1.0, 2.0, 3.0, np.inf))),
        draw(st.one_of(st.none(), st.floats(0.0, 4.0))),
        draw(st.sampled_from((1, 8, 16, 32, 64))),
    )


@settings(max_examples=60, deadline=None)
@given(problem=problems())
def test_oracle_compatibility_with_scipy(problem):
    data, queries, k, p, md, ls = problem
    d, i = OracleTree(data, ls).query(queries, k=k, p=p, max_distance=md)
    sd, si = SciPyKDTree(data, leafsize=ls).query(
        queries, k=k, p=p, distance_upper_bound=np.inf if md is None else md
    )
    shape = d.shape
    np.testing.assert_array_equal(i, np.asarray(si).reshape(shape))
    np.testing.assert_allclose(d, np.asarray(sd).reshape(shape), atol=1e-12)
