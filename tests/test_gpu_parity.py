"""GPU parity: the HIP build and query against the oracle, bit for bit.

Build parity follows the reference's own parallel_build_matches_serial_build
(src/build.rs:345-359): data, indices, nodes and boxes must be identical.  Query
parity compares distances and indices exactly (assert_array_equal), which is
stricter than the reference's SciPy check (tests/test_kdtree.py:88-89).
"""

from __future__ import annotations

import numpy as np
import pytest

from kdtree import KDTree
from oracle import OracleTree
from summation_cases import ORDER_CASES

pytestmark = pytest.mark.gpu


def hashed(n: int, d: int) -> np.ndarray:
    # The data of parallel_build_matches_serial_build, src/build.rs:348-350:
    # many repeated coordinates, so the pivot-equal fix-up runs.
    i = np.arange(n * d, dtype=np.uint64)
    return ((i * np.uint64(2_654_435_761)) % np.uint64(10_007)).astype(np.float64).reshape(n, d)


BUILD_CASES = {
    "uniform-1000x3-ls32": (lambda: np.random.default_rng(0).uniform(size=(1000, 3)), 32),
    "hashed-2000x5-ls8": (lambda: hashed(2000, 5), 8),
    "c1-10000x3-ls32": (lambda: np.random.default_rng(0).uniform(size=(10_000, 3)), 32),
    "normal-5000x8-ls16": (lambda: np.random.default_rng(3).normal(size=(5000, 8)), 16),
    "lattice-3000x2-ls1": (lambda: np.random.default_rng(4).integers(0, 7, size=(3000, 2)).astype(float), 1),
    "uniform-100000x3-ls32": (lambda: np.random.default_rng(5).uniform(size=(100_000, 3)), 32),
    "hashed-60000x4-ls6": (lambda: hashed(60_000, 4), 6),
    "normal-20000x16-ls64": (lambda: np.random.default_rng(6).normal(size=(20_000, 16)), 64),
    "line-513x1-ls6": (lambda: np.random.default_rng(7).normal(size=(513, 1)), 6),
    "single-point": (lambda: np.array([[1.5, -2.0]]), 32),
    "all-equal-300x3": (lambda: np.ones((300, 3)), 4),
    "signed-zeros": (lambda: np.array([[0.0], [-0.0]] * 50), 3),
    "wide-200x33-ls5": (lambda: np.random.default_rng(8).normal(size=(200, 33)), 5),
}


def assert_same_tree(gpu, ora):
    gd, gi, gn, gb = gpu
    od, oi, on, ob = ora
    np.testing.assert_array_equal(gi, oi)
    assert gd.tobytes() == od.tobytes(), "tree-order rows differ"
    assert len(gn) == len(on)
    for field in ("is_leaf", "a", "b", "split_dim", "order_by_box"):
        np.testing.assert_array_equal(gn[field], on[field], err_msg=field)
    inner = on["is_leaf"] == 0
    assert gn["split_value"][inner].tobytes() == on["split_value"][inner].tobytes()
    # Boxes: equal as numbers (a box may hold -0.0 where the other holds +0.0;
    # no query arithmetic can tell them apart).
    np.testing.assert_array_equal(gb, ob)


@pytest.mark.parametrize("case", BUILD_CASES.values(), ids=BUILD_CASES.keys())
def test_build_matches_oracle(case):
    make, leafsize = case
    data = make()
    print(f"build case n={data.shape[0]} d={data.shape[1]} leafsize={leafsize}", flush=True)
    tree = KDTree(data, leafsize=leafsize)
    assert_same_tree(tree._export(), OracleTree(data, leafsize).export())
    np.testing.assert_array_equal(tree.data, data)


def _query_case(data, queries, k, p=2.0, max_distance=None, eps=0.0, leafsize=32, fp32=False):
    print(f"query case n={data.shape} q={queries.shape} k={k} p={p} md={max_distance} eps={eps} "
          f"ls={leafsize} fp32={fp32}", flush=True)
    tree = KDTree(data, leafsize=leafsize)
    gd, gi = tree.query(queries, k=k, p=p, max_distance=max_distance, eps=eps, fp32_screen=fp32)
    od, oi = OracleTree(data, leafsize).query(queries, k=k, p=p, max_distance=max_distance, eps=eps)
    np.testing.assert_array_equal(gi, oi)
    if np.isfinite(p) and p not in (1.0, 2.0):
        # General p: device pow versus host libm, parity at 1e-12.
        np.testing.assert_allclose(gd, od, rtol=1e-12, atol=1e-12)
    else:
        assert gd.tobytes() == od.tobytes(), "distances differ in bits"


def test_c1_configuration():
    # n=10,000 d=3 U[0,1) seed 0, 1,000 queries seed 1, leafsize 32, k=8.
    data = np.random.default_rng(0).uniform(size=(10_000, 3))
    queries = np.random.default_rng(1).uniform(size=(1_000, 3))
    _query_case(data, queries, k=8)


@pytest.mark.parametrize("dims", list(range(1, 18)) + [24, 31, 32, 33, 63, 64])
@pytest.mark.parametrize("p", [1.0, 2.0, np.inf])
def test_row_widths(dims, p):
    # row-width-NN cases of tests/test_kdtree.py:153-163.
    data = np.random.default_rng(dims).normal(size=(512, dims))
    queries = np.random.default_rng(1_000 + dims).normal(size=(64, dims))
    _query_case(data, queries, k=3, p=p, leafsize=6)


@pytest.mark.parametrize("k", [1, 2, 7, 16, 17, 32, 33, 64, 65, 200])
def test_k_sizes(k):
    data = np.random.default_rng(11).normal(size=(3000, 3))
    queries = np.random.default_rng(12).normal(size=(200, 3))
    _query_case(data, queries, k=k, leafsize=16)


@pytest.mark.parametrize("leafsize", [1, 3, 8, 31, 32, 64, 100])
@pytest.mark.parametrize("dims", [4, 8])
def test_leafsizes_partial_chunks(leafsize, dims):
    # d4 / d8 leaf tails use a different summation order (src/kernel.rs:456).
    data = np.random.default_rng(20 + leafsize).normal(size=(2500, dims))
    queries = np.random.default_rng(21).normal(size=(100, dims))
    _query_case(data, queries, k=5, leafsize=leafsize)


@pytest.mark.parametrize("case", ORDER_CASES, ids=[c[0] for c in ORDER_CASES])
def test_summation_order_closed_forms(case):
    # The oracle's closed-form order vectors through the device kernels: one
    # leaf of `length` rows (length <= leafsize, so rows keep their input
    # order), the vector at row `offset`, every other row far away; L1, so the
    # returned distance is the sum of the terms itself (1 or 1 + 2^-52).
    _, d, v, off, length, want, _ = case
    data = 1000.0 + np.arange(length * d, dtype=np.float64).reshape(length, d)
    data[off] = v
    tree = KDTree(data, leafsize=32)
    assert tree._export()[1][off] == off
    gd, gi = tree.query(np.zeros((1, d)), k=length, p=1.0)
    assert gi[0, 0] == off
    assert gd[0, 0] == want, f"{gd[0, 0]!r} != {want!r}"


@pytest.mark.parametrize("p", [1.0, 2.0, 3.0, np.inf])
def test_general_metrics(p):
    data = np.random.default_rng(30).normal(size=(4000, 5))
    queries = np.random.default_rng(31).normal(size=(100, 5))
    _query_case(data, queries, k=6, p=p)


@pytest.mark.parametrize("max_distance", [0.0, 0.05, 0.3, 1.0])
def test_max_distance(max_distance):
    data = np.random.default_rng(40).normal(size=(3000, 3))
    queries = np.random.default_rng(41).normal(size=(150, 3))
    _query_case(data, queries, k=10, max_distance=max_distance)


@pytest.mark.parametrize("eps", [0.0, 0.1, 1.0])
def test_eps_visit_order(eps):
    # With eps > 0 the result depends on the visit order; the device walks the
    # tree in the reference's order, so it still matches.
    data = np.random.default_rng(50).normal(size=(5000, 3))
    queries = np.random.default_rng(51).normal(size=(200, 3))
    _query_case(data, queries, k=4, eps=eps)


def test_ties_lattice():
    # Integer lattice: many exactly equal distances; ties go to the smaller
    # original index (src/query.rs:415, :465, :474).
    g = np.stack(np.meshgrid(np.arange(12.0), np.arange(12.0), np.arange(6.0)), -1).reshape(-1, 3)
    data = np.concatenate([g, g[::7]])  # exact duplicates too
    queries = np.random.default_rng(60).integers(0, 12, size=(300, 3)).astype(float) + 0.5 * (
        np.random.default_rng(61).integers(0, 2, size=(300, 3))
    )
    _query_case(data, queries, k=9, leafsize=8)


def test_order_by_box_data():
    # Points on a 1-D curve in 3-D: boxes shrink off-axis, so order_by_box fires.
    t = np.random.default_rng(70).uniform(0, 10, size=6000)
    data = np.stack([t, np.sin(t), np.cos(t) * 0.01], 1)
    tree = KDTree(data, leafsize=8)
    assert tree._export()[2]["order_by_box"].sum() > 0
    queries = np.random.default_rng(71).uniform(-1, 11, size=(300, 3))
    _query_case(data, queries, k=5, leafsize=8)
    _query_case(data, queries, k=5, leafsize=8, eps=0.5)


def test_fp32_screen_is_exact():
    rng = np.random.default_rng(80)
    centres = rng.uniform(0, 10, size=(16, 8))
    data = centres[rng.integers(0, 16, size=20000)] + 0.3 * rng.normal(size=(20000, 8))
    queries = centres[rng.integers(0, 16, size=500)] + 0.3 * rng.normal(size=(500, 8))
    _query_case(data, queries, k=32, fp32=True)
    _query_case(data[:, :3].copy(), queries[:, :3].copy(), k=16, fp32=True)


def test_large_query_batch():
    data = np.random.default_rng(90).uniform(size=(200_000, 3))
    queries = np.random.default_rng(91).uniform(size=(20_000, 3))
    _query_case(data, queries, k=16)


@pytest.mark.parametrize("dims", [2, 3, 8, 16])
@pytest.mark.parametrize("k", [1, 8, 10, 16, 32])
def test_benchmark_shapes_match_oracle(dims, k):
    # The (ndim, k) pairs of the benchmark configurations.
    data = np.random.default_rng(100 + dims).normal(size=(6000, dims))
    queries = np.random.default_rng(200 + k).normal(size=(300, dims))
    _query_case(data, queries, k=k, leafsize=16)


def test_eps_box_order_and_cutoff_together():
    t = np.random.default_rng(110).uniform(0, 10, size=8000)
    data = np.stack([t, np.sin(t), 0.01 * np.cos(t)], 1)
    queries = np.random.default_rng(111).uniform(-1, 11, size=(400, 3))
    _query_case(data, queries, k=8, leafsize=8, eps=0.3)
    _query_case(data, queries, k=16, leafsize=8, max_distance=0.4)
    g = np.stack(np.meshgrid(np.arange(10.0), np.arange(10.0), np.arange(10.0)), -1).reshape(-1, 3)
    _query_case(np.concatenate([g, g[::3]]), g[::7] + 0.5, k=10, leafsize=4)


@pytest.mark.parametrize("no_order", [True, False])
def test_query_ordering_is_invisible(no_order):
    # Batches >= 8192 queries run in leaf order unless KDT_QUERY_NO_ORDER
    # (0x200) is set; results must not change.
    import ctypes

    data = np.random.default_rng(120).uniform(size=(50_000, 3))
    queries = np.random.default_rng(121).uniform(size=(20_000, 3))
    od, oi = OracleTree(data, 32).query(queries, k=16)
    tree = KDTree(data, leafsize=32)
    out_d = np.empty((queries.shape[0], 16))
    out_i = np.empty((queries.shape[0], 16), dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(queries.ctypes.data), queries.shape[0], 3, 16, 2.0,
                             np.inf, 0.0, 0x200 if no_order else 0, ctypes.c_void_p(out_d.ctypes.data),
                             ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    np.testing.assert_array_equal(out_i, oi)
    assert out_d.tobytes() == od.tobytes()


@pytest.mark.parametrize("group", [8, 16, 32, 64])
@pytest.mark.parametrize("dims,k", [(3, 1), (3, 8), (8, 8), (16, 5), (2, 16)])
def test_lanes_per_query_do_not_change_results(group, dims, k):
    # The group width (lanes per query, a per-tree autotuner choice) only
    # changes speed; widths below k fall back to the automatic one.
    import ctypes

    data = np.random.default_rng(130 + dims).normal(size=(9000, dims))
    queries = np.random.default_rng(140 + k).normal(size=(9000, dims))  # >= 8192: leaf-ordered path too
    od, oi = OracleTree(data, 16).query(queries, k=k)
    tree = KDTree(data, leafsize=16)
    code = {8: 1, 16: 2, 32: 3, 64: 4}[group]
    out_d = np.empty((queries.shape[0], k))
    out_i = np.empty((queries.shape[0], k), dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(queries.ctypes.data), queries.shape[0], dims, k, 2.0,
                             np.inf, 0.0, code << 4, ctypes.c_void_p(out_d.ctypes.data),
                             ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    np.testing.assert_array_equal(out_i, oi)
    assert out_d.tobytes() == od.tobytes()


WIDE_CASES = [(k, g, 2.0, 0.0, np.inf) for k, g in [(1, 0), (10, 0), (10, 32), (16, 64), (32, 0), (64, 0), (8, 8)]] + [
    (10, 0, 1.0, 0.0, np.inf), (10, 0, np.inf, 0.0, np.inf), (10, 0, 2.0, 0.5, np.inf), (10, 0, 2.0, 0.0, 9.0)]


@pytest.mark.parametrize("dims", [32, 64])
@pytest.mark.parametrize("k,group,p,eps,max_distance", WIDE_CASES)
def test_wide_rows_compiled(dims, k, group, p, eps, max_distance):
    # Rows of 32 and 64 coordinates run compiled register-list kernels (the
    # Streamed recipe, src/kernel.rs:177-188, :296-303); 8-lane requests take
    # 16 lanes.  9000 queries: the leaf-ordered path too.  Float32-valued
    # Gaussian data as in the serving workload.
    import ctypes

    rng = np.random.default_rng(300 + dims + k)
    data = rng.normal(size=(6000, dims)).astype(np.float32).astype(np.float64)
    queries = rng.normal(size=(9000, dims)).astype(np.float32).astype(np.float64)
    md = None if np.isinf(max_distance) else max_distance
    od, oi = OracleTree(data, 32).query(queries, k=k, p=p, eps=eps, max_distance=md, threads=8)
    tree = KDTree(data, leafsize=32)
    code = {0: 0, 8: 1, 16: 2, 32: 3, 64: 4}[group]
    out_d = np.empty((queries.shape[0], k))
    out_i = np.empty((queries.shape[0], k), dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(queries.ctypes.data), queries.shape[0], dims, k, p,
                             max_distance, eps, code << 4, ctypes.c_void_p(out_d.ctypes.data),
                             ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    np.testing.assert_array_equal(out_i, oi)
    assert out_d.tobytes() == od.tobytes()


@pytest.mark.parametrize("group", [8, 16, 32, 64])
@pytest.mark.parametrize("case", ["lattice", "tiny"])
def test_merge_ties_and_padding_every_group_width(group, case):
    # The k list's compare-exchange steps decide by one lexicographic test
    # (distance, then original index); exact ties, duplicate points and the
    # padding entries of a list with fewer than k points all reach it.  k equals
    # the group width, leaves of 32 rows give up to 32 hits per chunk, so both
    # the bulk merge and the single inserts run at every width.
    import ctypes

    g = np.stack(np.meshgrid(np.arange(10.0), np.arange(10.0), np.arange(5.0)), -1).reshape(-1, 3)
    data = np.concatenate([g, g[::3], g[::11]]) if case == "lattice" else g[: group // 2 + 3]
    rng = np.random.default_rng(200 + group)
    queries = rng.integers(0, 10, size=(2500, 3)).astype(float) + 0.5 * rng.integers(0, 2, size=(2500, 3))
    k = group
    od, oi = OracleTree(data, 32).query(queries, k=k)
    tree = KDTree(data, leafsize=32)
    code = {8: 1, 16: 2, 32: 3, 64: 4}[group]
    out_d = np.empty((queries.shape[0], k))
    out_i = np.empty((queries.shape[0], k), dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(queries.ctypes.data), queries.shape[0], 3, k, 2.0,
                             np.inf, 0.0, code << 4, ctypes.c_void_p(out_d.ctypes.data),
                             ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    if case == "tiny":
        assert (oi == data.shape[0]).any()  # padded rows are part of the check
    np.testing.assert_array_equal(out_i, oi)
    assert out_d.tobytes() == od.tobytes()
