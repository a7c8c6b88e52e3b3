"""Matrix-core screen for 32 <= ndim <= 64 (p = 2, eps = 0, k <= 32).

Every result must equal the oracle's (the reference's descent restated) bit
for bit, whichever way a query is answered: by the screen + FP64 recheck +
visit certificate, or by the traversal for the queries the certificate does
not cover.  The screen-only diagnostic flag skips that traversal, so the rows
it writes show what the matrix-core path itself answered.
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
MATRIX_SCREEN = 0x8


def f32_normal(seed: int, n: int, d: int) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32).astype(np.float64)


def mixture(seed: int, n: int, d: int, sigma: float) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 10, size=(16, d))
    return centres[rng.integers(0, 16, size=n)] + sigma * rng.standard_normal((n, d))


# name: (data, queries, k, leafsize, max_distance, (min, max) fraction of the
# queries the screen answers by itself).  Duplicated rows overflow the
# candidate list, so every such query must take the traversal.
CASES = {
    "normal-f32-20000x64-k10": (lambda: f32_normal(1, 20_000, 64), lambda: f32_normal(2, 1000, 64), 10, 32, None, (0.95, 1.0)),
    "normal-f32-6000x32-k16": (lambda: f32_normal(3, 6000, 32), lambda: f32_normal(4, 777, 32), 16, 32, None, (0.95, 1.0)),
    "normal-3001x33-k1": (lambda: np.random.default_rng(5).normal(size=(3001, 33)),
                          lambda: np.random.default_rng(6).normal(size=(65, 33)), 1, 32, None, (0.95, 1.0)),
    "normal-777x48-k16-ls8": (lambda: np.random.default_rng(7).normal(size=(777, 48)),
                              lambda: np.random.default_rng(8).normal(size=(300, 48)), 16, 8, None, (0.95, 1.0)),
    "uniform-f64-10000x40-k5": (lambda: np.random.default_rng(9).uniform(size=(10_000, 40)),
                                lambda: np.random.default_rng(10).uniform(size=(500, 40)), 5, 32, None, (0.95, 1.0)),
    "mixture-8000x64-k10": (lambda: mixture(11, 8000, 64, 0.3), lambda: mixture(12, 400, 64, 0.3), 10, 16, None, (0.0, 1.0)),
    "duplicates-4000x64-k10": (lambda: np.repeat(f32_normal(13, 40, 64), 100, axis=0),
                               lambda: f32_normal(14, 100, 64), 10, 32, None, (0.0, 0.0)),
    "tiny-5x64-k10": (lambda: f32_normal(15, 5, 64), lambda: f32_normal(16, 33, 64), 10, 32, None, (0.95, 1.0)),
    "cutoff-5000x64-k16": (lambda: f32_normal(17, 5000, 64), lambda: f32_normal(18, 200, 64), 16, 32, 10.0, (0.95, 1.0)),
    "far-queries-5000x32-k4": (lambda: f32_normal(19, 5000, 32), lambda: 50.0 + f32_normal(20, 100, 32), 4, 32,
                               None, (0.0, 1.0)),
    # k above the 16-entry lane lists: more point slices per query
    "normal-f32-20000x64-k32": (lambda: f32_normal(31, 20_000, 64), lambda: f32_normal(32, 600, 64), 32, 32, None,
                                (0.9, 1.0)),
    "normal-6000x40-k20-cutoff": (lambda: np.random.default_rng(33).normal(size=(6000, 40)),
                                  lambda: np.random.default_rng(34).normal(size=(300, 40)), 20, 16, 9.0, (0.0, 1.0)),
}


def query_flags(tree: KDTree, q: np.ndarray, k: int, md, flags: int):
    out_d = np.full((q.shape[0], k), np.nan)
    out_i = np.full((q.shape[0], k), -7, dtype=np.int64)
    rc = tree._lib.kdt_query(tree._h, ctypes.c_void_p(q.ctypes.data), q.shape[0], q.shape[1], k, 2.0,
                             np.inf if md is None else md, 0.0, flags,
                             ctypes.c_void_p(out_d.ctypes.data), ctypes.c_void_p(out_i.ctypes.data))
    assert rc == 0, tree._lib.kdt_last_error().decode()
    return out_d, out_i


def screen_only(tree: KDTree, q: np.ndarray, k: int, md, extra: int = 0):
    return query_flags(tree, q, k, md, FP32_SCREEN | SCREEN_ONLY | extra)


@pytest.mark.parametrize("case", CASES.values(), ids=CASES.keys())
def test_screen_matches_oracle(case):
    make_data, make_q, k, leafsize, md, (lo, hi) = case
    data = np.ascontiguousarray(make_data())
    q = np.ascontiguousarray(make_q())
    od, oi = OracleTree(data, leafsize).query(q, k=k, max_distance=md)
    tree = KDTree(data, leafsize=leafsize)
    gd, gi = tree.query(q, k=k, max_distance=md, fp32_screen=True)
    np.testing.assert_array_equal(gi, oi)
    assert gd.tobytes() == od.tobytes()
    # The plain FP64 path agrees too.
    pd, pi = tree.query(q, k=k, max_distance=md)
    np.testing.assert_array_equal(pi, oi)
    assert pd.tobytes() == od.tobytes()
    # Rows the matrix-core path answered by itself are exact as well.
    sd, si = screen_only(tree, q, k, md)
    done = si[:, 0] != -7
    np.testing.assert_array_equal(si[done], oi[done])
    assert sd[done].tobytes() == od[done].tobytes()
    print(f"screen answered {done.mean():.3f} of the queries", flush=True)
    assert lo <= done.mean() <= hi


NARROW = {
    "normal-f32-20000x16-k10": (lambda: f32_normal(23, 20_000, 16), lambda: f32_normal(24, 700, 16), 10, 32),
    "uniform-9000x16-k16-ls8": (lambda: np.random.default_rng(25).uniform(size=(9000, 16)),
                                lambda: np.random.default_rng(26).uniform(size=(300, 16)), 16, 8),
    "normal-4000x24-k3": (lambda: np.random.default_rng(27).normal(size=(4000, 24)),
                          lambda: np.random.default_rng(28).normal(size=(129, 24)), 3, 32),
    "normal-f32-12000x16-k32": (lambda: f32_normal(35, 12_000, 16), lambda: f32_normal(36, 500, 16), 32, 32),
    # 8 <= ndim < 16: rows padded to 16 columns
    "uniform-f32-20000x8-k8": (lambda: np.random.default_rng(39).uniform(size=(20_000, 8)).astype(np.float32)
                               .astype(np.float64),
                               lambda: np.random.default_rng(40).uniform(size=(600, 8)).astype(np.float32)
                               .astype(np.float64), 8, 32),
    "normal-9000x12-k32-ls16": (lambda: np.random.default_rng(41).normal(size=(9000, 12)),
                                lambda: np.random.default_rng(42).normal(size=(400, 12)), 32, 16),
    "normal-5000x8-k1": (lambda: np.random.default_rng(43).normal(size=(5000, 8)),
                         lambda: np.random.default_rng(44).normal(size=(257, 8)), 1, 32),
}


@pytest.mark.parametrize("case", NARROW.values(), ids=NARROW.keys())
def test_matrix_screen_narrow_rows(case):
    """8 <= ndim < 32 takes the matrix-core screen only when asked."""
    make_data, make_q, k, leafsize = case
    data = np.ascontiguousarray(make_data())
    q = np.ascontiguousarray(make_q())
    od, oi = OracleTree(data, leafsize).query(q, k=k)
    tree = KDTree(data, leafsize=leafsize)
    gd, gi = query_flags(tree, q, k, None, FP32_SCREEN | MATRIX_SCREEN)
    np.testing.assert_array_equal(gi, oi)
    assert gd.tobytes() == od.tobytes()
    sd, si = screen_only(tree, q, k, None, MATRIX_SCREEN)
    done = si[:, 0] != -7
    np.testing.assert_array_equal(si[done], oi[done])
    assert sd[done].tobytes() == od[done].tobytes()
    assert done.mean() >= 0.95
    # Without the request the traversal answers (nothing left unwritten).
    _, ti = screen_only(tree, q, k, None)
    np.testing.assert_array_equal(ti, oi)


def test_k32_uniform_rows_stay_on_the_screen():
    # Rows in tree order put a query's near neighbours in one stretch of point
    # tiles.  For k > 16 the point slices interleave their tiles; as contiguous
    # ranges, the two lane lists of the slice holding that stretch overflowed
    # and 12% of these queries were left to the (much slower) traversal.
    def f32_uniform(seed, n):
        return np.random.default_rng(seed).uniform(size=(n, 16)).astype(np.float32).astype(np.float64)

    data, q = f32_uniform(37, 100_000), f32_uniform(38, 20_000)
    tree = KDTree(data, leafsize=32)
    pd, pi = tree.query(q, k=32)  # the FP64 traversal
    sd, si = screen_only(tree, q, 32, None, MATRIX_SCREEN)
    done = si[:, 0] != -7
    print(f"screen answered {done.mean():.4f} of the queries", flush=True)
    assert done.mean() >= 0.99
    np.testing.assert_array_equal(si[done], pi[done])
    assert sd[done].tobytes() == pd[done].tobytes()
    od, oi = OracleTree(data, 32).query(q[:300], k=32, threads=8)
    np.testing.assert_array_equal(pi[:300], oi)
    assert pd[:300].tobytes() == od.tobytes()


@pytest.mark.parametrize("k", [10, 32])
def test_screen_large_batch_device(k):
    """2^14 queries on the device: few point slices per query tile (k = 32
    takes at least 8, so that a query's top k spreads over its lists)."""
    import torch

    from kdtree import forest as F

    data = f32_normal(21, 50_000, 64)
    q = f32_normal(22, 1 << 14, 64)
    od, oi = OracleTree(data, 32).query(q, k=k, threads=8)
    tree = F.build(data, 32)
    try:
        qd = torch.from_numpy(q).cuda()
        gd, gi = F.query_device(tree, qd, k, flags=F.KDT_QUERY_ASSUME_FINITE | F.KDT_QUERY_FP32_SCREEN)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(gi.cpu().numpy(), oi)
        assert gd.cpu().numpy().tobytes() == od.tobytes()
    finally:
        tree.free()


@pytest.mark.parametrize("d,k", [(64, 10), (16, 16)])
def test_seeded_clusters_shared_threshold(d, k):
    """A seeded tree (>= 1024 point tiles) of tight clusters: many lists of a
    query fill up with values under tau_A, and every list of a query lowers
    the threshold of all the others (a value at or above any of its lists' last
    entries is dropped).  Queries whose lists cannot certify completeness must
    be flagged and answered by the traversal; every answer equals the
    oracle's bit for bit, and some queries are answered by the screen itself."""
    import torch

    from kdtree import forest as F

    rng = np.random.default_rng(61 + d)
    centres = rng.uniform(-5, 5, size=(500, d))
    scale = np.where(np.arange(500) < 250, 0.02, 0.3)[:, None]  # tight clusters: whole lists under tau_A
    lab = rng.integers(0, 500, size=48_000)
    data = (centres[lab] + scale[lab] * rng.standard_normal((48_000, d))).astype(np.float32).astype(np.float64)
    ql = rng.integers(0, 500, size=1500)
    q = np.concatenate([centres[ql] + scale[ql] * rng.standard_normal((1500, d)),
                        rng.uniform(-5, 5, size=(500, d))]).astype(np.float32).astype(np.float64)
    od, oi = OracleTree(data, 32).query(q, k=k, threads=8)
    tree = F.build(data, 32)
    try:
        qd = torch.from_numpy(q).cuda()
        flags = F.KDT_QUERY_ASSUME_FINITE | F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_MATRIX_SCREEN
        gd, gi = F.query_device(tree, qd, k, flags=flags)
        sd = torch.full_like(gd, -7.0)
        si = torch.full_like(gi, -7)
        F.query_device(tree, qd, k, sd, si, flags=flags | F.KDT_QUERY_SCREEN_ONLY)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(gi.cpu().numpy(), oi)
        assert gd.cpu().numpy().tobytes() == od.tobytes()
        done = (si[:, 0] != -7).cpu().numpy()
        print(f"{d}-d clusters, k = {k}: screen answered {done.mean():.3f}", flush=True)
        assert 0.2 < done.mean() < 1.0
        np.testing.assert_array_equal(si.cpu().numpy()[done], oi[done])
    finally:
        tree.free()


@pytest.mark.parametrize("d,k", [(64, 10), (16, 32)])
def test_screen_slice_shapes_same_results(d, k):
    """KDT_QUERY_SLICES_*: the matrix-core screen's point slices at the
    occupancy rule, half, twice and four times it (the per-tree work-tile
    shape the autotuner picks) all give the oracle's answer bit for bit."""
    import torch

    from kdtree import forest as F

    data = f32_normal(91 + d, 40_000, d)  # 1,250 point tiles: seeded
    q = f32_normal(92 + d, 3000, d)
    od, oi = OracleTree(data, 32).query(q, k=k, threads=8)
    tree = F.build(data, 32)
    try:
        qd = torch.from_numpy(q).cuda()
        for code in range(4):
            flags = (F.KDT_QUERY_ASSUME_FINITE | F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_MATRIX_SCREEN
                     | (code << F.KDT_QUERY_SLICES_SHIFT))
            gd, gi = F.query_device(tree, qd, k, flags=flags)
            torch.cuda.synchronize()
            print(f"{d}-d k = {k}, slice code {code}", flush=True)
            np.testing.assert_array_equal(gi.cpu().numpy(), oi)
            assert gd.cpu().numpy().tobytes() == od.tobytes()
    finally:
        tree.free()
