"""The serving loop: requests that overlap on the device get exactly the result
they get alone (through the plain FP64 kernel), and the oracle's."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import OracleTree

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone_and_oracle():
    import torch

    from kdtree import forest as F
    from kdtree.serve import Server, latency_stats, poisson_requests

    dims = [16, 64, 16, 64]
    rng = np.random.default_rng(7)
    datas = [rng.standard_normal((int(n), d)).astype(np.float32).astype(np.float64)
             for n, d in zip([3000, 1500, 5000, 2500], dims)]
    trees = [F.build(x, 32) for x in datas]

    def make_q(i, m, d):
        return np.random.default_rng(1000 + i).standard_normal((m, d)).astype(np.float32).astype(np.float64)

    reqs = poisson_requests(40, rate=2000.0, tree_dims=dims, k=10, max_batch=400, seed=3, make_queries=make_q)
    # Served with the FP32 screens (matrix-core screen for every tree here),
    # requests above 96 queries in slices; compared below with the plain FP64
    # traversal of each request alone.
    server = Server(trees, n_streams=3, max_queries_per_launch=600, flags=F.KDT_QUERY_FP32_SCREEN | 0x8,
                    slice_max=96)
    server.run(reqs, time_limit=120)
    stats = latency_stats(reqs)
    print(stats, "max_inflight", server.max_inflight, "launches", server.launches, flush=True)
    assert server.launches < len(reqs)  # requests were coalesced
    assert server.max_inflight >= 2  # launches overlapped on the device
    for r in reqs:
        q = torch.from_numpy(r.queries).cuda()
        ad, ai = F.query_device(trees[r.tree], q, r.k, flags=0)
        np.testing.assert_array_equal(ai.cpu().numpy(), r.out_i)
        assert ad.cpu().numpy().tobytes() == r.out_d.tobytes()
    for r in reqs[:6]:
        od, oi = OracleTree(datas[r.tree], 32).query(r.queries, k=r.k)
        np.testing.assert_array_equal(oi, r.out_i)
        assert od.tobytes() == r.out_d.tobytes()
    for t in trees:
        t.free()


def test_small_request_behind_a_large_one_finishes_first():
    # A large request (sliced) arrives first, a one-query request just after
    # it: the small-request lane and the slicing let the small one finish long
    # before the large one, and both still get their exact answers.
    import torch

    from kdtree import forest as F
    from kdtree.serve import Request, Server

    rng = np.random.default_rng(11)
    data = rng.standard_normal((200_000, 64)).astype(np.float32).astype(np.float64)
    tree = F.build(data, 32)
    try:
        big = Request(0, 0.0, 0, rng.standard_normal((65_536, 64)).astype(np.float32).astype(np.float64), 10)
        small = Request(1, 0.005, 0, rng.standard_normal((1, 64)).astype(np.float32).astype(np.float64), 10)
        server = Server([tree], n_streams=4, flags=F.KDT_QUERY_FP32_SCREEN)
        server.run([big, small], time_limit=120)
        print({"big_ms": big.done * 1e3, "small_ms": (small.done - small.arrival) * 1e3,
               "launches": server.launches, "max_inflight": server.max_inflight}, flush=True)
        assert small.done < big.done
        for r in (big, small):
            q = torch.from_numpy(r.queries).cuda()
            ad, ai = F.query_device(tree, q, r.k, flags=F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_ASSUME_FINITE)
            np.testing.assert_array_equal(ai.cpu().numpy(), r.out_i)
            assert ad.cpu().numpy().tobytes() == r.out_d.tobytes()
    finally:
        torch.cuda.synchronize()
        tree.free()
