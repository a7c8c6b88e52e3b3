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


def test_non_finite_queries_leave_outputs_untouched(small_tree):
    # The finiteness check runs before the query (src/query.rs:46): a call that
    # fails has written nothing into the caller's buffers, through the host
    # entry point and the device one alike.
    import ctypes

    import torch

    from kdtree._core import load_library

    lib = load_library()
    q = np.array([[0.1, 0.2], [np.nan, 0.0], [0.3, 0.3]])
    out_d = np.full((3, 2), 7.0)
    out_i = np.full((3, 2), 7, np.int64)
    rc = lib.kdt_query(small_tree._handle, ctypes.c_void_p(q.ctypes.data), 3, 2, 2, 2.0, float("inf"), 0.0, 0,
                       ctypes.c_void_p(out_d.ctypes.data), ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 7 and lib.kdt_last_error().decode() == "all coordinates must be finite"
    assert np.all(out_d == 7.0) and np.all(out_i == 7)
    qd = torch.from_numpy(q).cuda()
    dd = torch.full((3, 2), 7.0, dtype=torch.float64, device="cuda")
    di = torch.full((3, 2), 7, dtype=torch.int64, device="cuda")
    rc = lib.kdt_query_device(small_tree._handle, ctypes.c_void_p(qd.data_ptr()), 3, 2, 2, 2.0, float("inf"), 0.0,
                              0, ctypes.c_void_p(dd.data_ptr()), ctypes.c_void_p(di.data_ptr()),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 7
    torch.cuda.synchronize()
    assert bool(torch.all(dd == 7.0)) and bool(torch.all(di == 7))


def test_batch_item_overflow_checks(small_tree):
    # The batch entry point checks each item like kdt_query_device: k beyond
    # 32 bits and n_queries * k overflow are rejected before any launch.
    import torch

    from kdtree import forest as F

    lib = F._lib()
    q = torch.zeros((4, 2), dtype=torch.float64, device="cuda")
    out = torch.zeros((4, 2), dtype=torch.float64, device="cuda")
    idx = torch.zeros((4, 2), dtype=torch.int64, device="cuda")
    for nq, k in ((4, 1 << 32), (1 << 62, 8)):
        items = (F.BatchItem * 1)()
        items[0] = F.BatchItem(small_tree._handle.value, q.data_ptr(), nq, 2, k, out.data_ptr(), idx.data_ptr(), 0, 0)
        rc = lib.kdt_query_batch_device(items, 1, 2.0, float("inf"), 0.0, 0, None)
        assert rc == 101 and "overflows" in lib.kdt_last_error().decode()
