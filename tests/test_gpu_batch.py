"""Multi-tree batches: one persistent launch, per-tree launches and a replayed
HIP graph all give, for every item, the oracle's answer for that tree (built
with the autotuned leafsize)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import OracleTree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def forest():
    import torch

    from kdtree import forest as F
    from kdtree.workloads import c4_specs, c4_tree_inputs

    specs = c4_specs(n_trees=24, scale=0.03)
    items, host = [], []
    for t, (n, d, q, k) in enumerate(specs):
        data, queries = c4_tree_inputs(t, n, d, q)
        print(f"tree {t}: n={n} d={d} q={q} k={k}", flush=True)
        tree = F.autotune_build(data, queries[:512], k, reps=1)
        items.append((tree, torch.from_numpy(queries).cuda(), k))
        host.append((data, queries, k, tree.leafsize))
    batch = F.Batch(items)
    yield batch, host
    batch.close()
    for tree, _, _ in items:
        tree.free()


def check_all(batch, host):
    for j, (data, queries, k, ls) in enumerate(host):
        od, oi = OracleTree(data, ls).query(queries, k=k)
        gd, gi = batch.results(j)
        np.testing.assert_array_equal(gi.cpu().numpy(), oi, err_msg=f"item {j}")
        assert gd.cpu().numpy().tobytes() == od.tobytes(), f"item {j} distances"


@pytest.mark.parametrize("mode", ["persistent", "per_tree", "plan"])
def test_batch_modes_match_oracle(forest, mode):
    import torch

    batch, host = forest
    batch.run(mode)
    torch.cuda.synchronize()
    check_all(batch, host)


def test_graph_replays_match_oracle(forest):
    import torch

    batch, host = forest
    for replay in range(3):
        for _, _, _, od, oi in batch.items:  # eager work between replays
            od.fill_(-1.0)
            oi.fill_(-1)
        torch.cuda.synchronize()
        batch.run("graph")
        torch.cuda.synchronize()
        check_all(batch, host)


def test_no_order_call_captures_as_a_pure_enqueue(forest):
    # kdt_query_batch_device with KDT_BATCH_NO_ORDER alone (no per-tree flag):
    # every item on its own launch sequence on the caller's stream, nothing
    # allocated or copied from the host, so a captured graph replays it.
    import torch

    from kdtree import forest as F

    batch, host = forest
    lib = F._lib()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        assert lib.kdt_query_batch_device(batch.arr, len(batch.items), 2.0, float("inf"), 0.0,
                                          F.KDT_BATCH_NO_ORDER, F._stream_ptr(s)) == 0
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        assert lib.kdt_query_batch_device(batch.arr, len(batch.items), 2.0, float("inf"), 0.0,
                                          F.KDT_BATCH_NO_ORDER, F._stream_ptr(None)) == 0
    for replay in range(3):
        for _, _, _, od, oi in batch.items:
            od.fill_(-1.0)
            oi.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check_all(batch, host)


def test_batch_fp32_screen_matches(forest):
    import torch

    batch, host = forest
    batch.run("persistent", extra_flags=0x1)
    torch.cuda.synchronize()
    check_all(batch, host)


def test_item_screen_flags_match_oracle(forest):
    # Items whose tree asks for the matrix-core screen (the per-item flags an
    # autotuner sets) run on their own launch sequence; every mode, the graph
    # replays included, still gives the oracle's answer for every item.
    import torch

    from kdtree import forest as F

    batch, host = forest
    sflags = F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_MATRIX_SCREEN
    trees = [t for t, _, _, _, _ in batch.items]
    saved = [t.flags for t in trees]
    forced = 0
    for tree, _, k, _, _ in batch.items:
        if 16 <= tree.d <= 64 and k <= 32:
            tree.flags = sflags
            forced += 1
    assert forced > 0, "no item the matrix-core screen applies to"
    try:
        b2 = F.Batch([(t, q, k) for t, q, k, _, _ in batch.items])
        for mode in ("persistent", "per_tree", "plan", "graph", "graph", "graph"):
            for _, _, _, od, oi in b2.items:
                od.fill_(-1.0)
                oi.fill_(-1)
            torch.cuda.synchronize()
            b2.run(mode)
            torch.cuda.synchronize()
            check_all(b2, host)
        b2.close()
    finally:
        for t, f in zip(trees, saved):
            t.flags = f


@pytest.mark.parametrize("p,eps,max_distance", [(1.0, 0.0, np.inf), (np.inf, 0.0, np.inf), (2.0, 0.5, np.inf),
                                                (2.0, 0.0, 0.3), (1.0, 0.25, 0.5)])
@pytest.mark.parametrize("per_tree", [False, True])
def test_batch_metrics_eps_cutoff_match_oracle(p, eps, max_distance, per_tree):
    # The batch call's p, eps and max_distance apply to every item, through
    # the shared launches and through per-tree launches alike.
    import ctypes

    import torch

    from kdtree import forest as F

    rng = np.random.default_rng(300)
    specs = [(3000, 2, 700, 8), (5000, 3, 900, 1), (2000, 8, 400, 32), (1500, 16, 300, 8), (40, 3, 50, 16),
             (9000, 3, 9000, 4)]
    items, host = [], []
    for n, d, q, k in specs:
        data = rng.uniform(size=(n, d))
        queries = rng.uniform(size=(q, d))
        tree = F.build(torch.from_numpy(data).cuda(), 16)
        items.append((tree, torch.from_numpy(queries).cuda(), k))
        host.append((data, queries, k))
    batch = F.Batch(items)
    lib = F._lib()
    flags = F.KDT_BATCH_PER_TREE_LAUNCHES if per_tree else 0
    rc = lib.kdt_query_batch_device(batch.arr, len(batch.items), ctypes.c_double(p), ctypes.c_double(max_distance),
                                    ctypes.c_double(eps), ctypes.c_uint32(flags), None)
    assert rc == 0, lib.kdt_last_error().decode()
    torch.cuda.synchronize()
    try:
        for j, (data, queries, k) in enumerate(host):
            od, oi = OracleTree(data, 16).query(queries, k=k, p=p, eps=eps,
                                                max_distance=None if np.isinf(max_distance) else max_distance)
            gd, gi = batch.results(j)
            np.testing.assert_array_equal(gi.cpu().numpy(), oi, err_msg=f"item {j}")
            assert gd.cpu().numpy().tobytes() == od.tobytes(), f"item {j} distances"
    finally:
        for tree, _, _ in items:
            tree.free()


def test_side_streams_join_into_caller_stream():
    # Items of 8192+ queries run their own launch sequences, spread over the
    # caller's stream and side streams forked from it; waiting on the caller's
    # stream alone (no device-wide synchronisation) must see every result.
    import torch

    from kdtree import forest as F

    rng = np.random.default_rng(90)
    items, host = [], []
    for t in range(6):
        data = rng.uniform(size=(40_000, 3))
        q = rng.uniform(size=(12_000 + 1_000 * t, 3))
        items.append((F.build(data, 32), torch.from_numpy(q).cuda(), 8))
        host.append((data, q))
    batch = F.Batch(items)
    want = [OracleTree(data, 32).query(q, k=8, threads=8) for data, q in host]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    try:
        for mode in ("persistent", "per_tree", "plan"):
            print(f"mode {mode}", flush=True)
            with torch.cuda.stream(s):
                for _, _, _, od, oi in batch.items:
                    od.fill_(-1.0)
                    oi.fill_(-1)
                batch.run(mode, stream=s)
                got = []
                for _, _, _, od, oi in batch.items:
                    hd = torch.empty(od.shape, dtype=od.dtype, pin_memory=True)
                    hi = torch.empty(oi.shape, dtype=oi.dtype, pin_memory=True)
                    hd.copy_(od, non_blocking=True)
                    hi.copy_(oi, non_blocking=True)
                    got.append((hd, hi))
            s.synchronize()
            for j, ((hd, hi), (od, oi)) in enumerate(zip(got, want)):
                np.testing.assert_array_equal(hi.numpy(), oi, err_msg=f"{mode} item {j}")
                assert hd.numpy().tobytes() == od.tobytes(), f"{mode} item {j} distances"
    finally:
        torch.cuda.synchronize()
        batch.close()
        for tree, _, _ in items:
            tree.free()
