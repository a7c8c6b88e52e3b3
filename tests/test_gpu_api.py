"""The Python surface of kdtree.KDTree on the GPU: the reference's VALID and
INVALID cases (tests/test_kdtree.py:98-180), its check_query assertions
(:76-89) against the committed SciPy golden vectors, and the error texts of
src/error.rs:18-51."""

from __future__ import annotations

import sys
import threading
from pathlib import Path

import numpy as np
import pytest

from kdtree import KDTree

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
VALID = make_golden.valid_cases()


@pytest.mark.parametrize("case", VALID, ids=[c[0] for c in VALID])
def test_valid(case):
    name, data, queries, k, p, md, ls = case
    print(f"valid case {name}", flush=True)
    g = np.load(GOLDEN / "valid.npz")
    tree = KDTree(data, leafsize=ls)
    d, i = tree.query(queries, k=k, p=p, max_distance=md)
    assert tree.ndim == data.shape[1]
    assert tree.n_points == data.shape[0]
    assert tree.leafsize == ls
    assert len(tree) == data.shape[0]
    np.testing.assert_array_equal(tree.data, data)
    shape = (k,) if np.ndim(queries) == 1 else (np.shape(queries)[0], k)
    assert d.shape == shape and i.shape == shape
    assert d.dtype == np.float64 and i.dtype == np.int64
    np.testing.assert_allclose(d, g[f"{name}/dist"], atol=1e-12)
    np.testing.assert_array_equal(i, g[f"{name}/idx"])


def test_c1_against_scipy_golden():
    data, queries = make_golden.c1_inputs()
    g = np.load(GOLDEN / "c1.npz")
    d, i = KDTree(data, leafsize=32).query(queries, k=8)
    np.testing.assert_array_equal(i, g["idx"])
    np.testing.assert_allclose(d, g["dist"], atol=1e-12)


@pytest.mark.parametrize(
    "data,queries,k",
    [
        (np.array([[0.0, 0.0], [1.0, 0.0]]), np.array([0.0]), 1),  # query-wrong-dimensions
        (np.array([[0.0, 0.0], [1.0, 0.0]]), np.array([0.0, 0.0]), 0),  # zero-k
    ],
    ids=["query-wrong-dimensions", "zero-k"],
)
def test_invalid(data, queries, k):
    with pytest.raises(ValueError):
        KDTree(data).query(queries, k=k)


@pytest.fixture(scope="module")
def small_tree():
    return KDTree(np.array([[0.0, 0.0], [1.0, 0.0], [4.0, 0.0]]))


@pytest.mark.parametrize(
    "kwargs,message",
    [
        ({"k": 0}, "k must be greater than zero"),
        ({"eps": -1.0}, "eps must be finite and non-negative, got -1"),
        ({"eps": float("nan")}, "eps must be finite and non-negative, got NaN"),
        ({"eps": float("inf")}, "eps must be finite and non-negative, got inf"),
        ({"max_distance": -0.5}, "max_distance must be finite and non-negative, or infinity, got -0.5"),
        ({"max_distance": float("nan")}, "max_distance must be finite and non-negative, or infinity, got NaN"),
        ({"p": 0.5}, "p must be finite and >= 1, or infinity, got 0.5"),
        ({"p": float("-inf")}, "p must be finite and >= 1, or infinity, got -inf"),
        ({"p": 1e-7}, "p must be finite and >= 1, or infinity, got 0.0000001"),
        ({"p": -1e21}, "p must be finite and >= 1, or infinity, got -1000000000000000000000"),
        ({"p": -0.0}, "p must be finite and >= 1, or infinity, got -0"),
    ],
)
def test_query_error_texts(small_tree, kwargs, message):
    with pytest.raises(ValueError) as e:
        small_tree.query(np.array([0.0, 0.0]), **kwargs)
    assert str(e.value) == message


def test_query_error_order(small_tree):
    # dimension first (src/lib.rs:173-174), then k, shape, finiteness, eps, ...
    with pytest.raises(ValueError, match="^dimension mismatch: expected 2, got 3$"):
        small_tree.query(np.zeros(3), k=0)
    with pytest.raises(ValueError, match="^k must be greater than zero$"):
        small_tree.query(np.zeros((0, 2)), k=0)
    with pytest.raises(ValueError, match="^queries must be a contiguous row-major matrix$"):
        small_tree.query(np.zeros((0, 2)))
    with pytest.raises(ValueError, match="^all coordinates must be finite$"):
        small_tree.query(np.array([np.inf, 0.0]), eps=-1.0)
    with pytest.raises(ValueError, match="^eps must"):
        small_tree.query(np.array([0.0, 0.0]), eps=-1.0, p=0.5)
    with pytest.raises(ValueError, match="^query must be one- or two-dimensional$"):
        small_tree.query(np.zeros((1, 1, 2)))


def test_build_errors():
    with pytest.raises(ValueError, match="^all coordinates must be finite$"):
        KDTree(np.array([[0.0, np.nan]]))
    with pytest.raises(TypeError, match="expected a numpy.ndarray, got list"):
        KDTree(np.zeros((2, 2))).query([0.0, 0.0])


def test_repr_len_and_data_copy():
    data = np.random.default_rng(3).normal(size=(100, 4))
    tree = KDTree(data, leafsize=8)
    assert repr(tree) == "KDTree(n_points=100, ndim=4, leafsize=8)"
    assert len(tree) == 100
    data[0, 0] = 1e9  # the tree owns a copy (src/lib.rs:116)
    assert tree.data[0, 0] != 1e9
    with pytest.raises(AttributeError):
        tree.leafsize = 3


def test_concurrent_queries_on_one_tree():
    data = np.random.default_rng(4).normal(size=(20000, 3))
    tree = KDTree(data)
    qs = [np.random.default_rng(10 + t).normal(size=(3000, 3)) for t in range(6)]
    want = [tree.query(q, k=5) for q in qs]
    got = [None] * len(qs)

    def run(t):
        got[t] = tree.query(qs[t], k=5)

    threads = [threading.Thread(target=run, args=(t,)) for t in range(len(qs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for (wd, wi), (gd, gi) in zip(want, got):
        np.testing.assert_array_equal(wi, gi)
        np.testing.assert_array_equal(wd, gd)


def test_float32_and_int_inputs_promote():
    data32 = np.random.default_rng(5).uniform(size=(500, 3)).astype(np.float32)
    tree = KDTree(data32)
    np.testing.assert_array_equal(tree.data, data32.astype(np.float64))
    d, i = tree.query(data32[:10], k=1)
    np.testing.assert_array_equal(i[:, 0], np.arange(10))
    np.testing.assert_array_equal(d[:, 0], np.zeros(10))


def test_large_host_results_staged():
    # 2^18 queries x k = 32: 128 MiB of results, copied to the caller's
    # (pageable) arrays through the library's pinned staging by several host
    # threads.  Every byte must equal the device-resident path's, and a sample
    # the oracle's; two calls at once on other threads must too.
    import torch

    from kdtree import forest as F
    from oracle import OracleTree

    rng = np.random.default_rng(95)
    data = rng.uniform(size=(200_000, 3))
    q = rng.uniform(size=(1 << 18, 3))
    k = 32
    tree = KDTree(data, leafsize=32)
    hd, hi = tree.query(q, k=k)
    dtree = F.build(data, 32)
    try:
        dd, di = F.query_device(dtree, torch.from_numpy(q).cuda(), k)
        torch.cuda.synchronize()
        assert hd.tobytes() == dd.cpu().numpy().tobytes()
        np.testing.assert_array_equal(hi, di.cpu().numpy())
    finally:
        dtree.free()
    od, oi = OracleTree(data, 32).query(q[:3000], k=k, threads=8)
    np.testing.assert_array_equal(hi[:3000], oi)
    assert hd[:3000].tobytes() == od.tobytes()
    out = {}
    threads = [threading.Thread(target=lambda j=j: out.setdefault(j, tree.query(q, k=k))) for j in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for j in range(2):
        assert out[j][0].tobytes() == hd.tobytes() and np.array_equal(out[j][1], hi)
