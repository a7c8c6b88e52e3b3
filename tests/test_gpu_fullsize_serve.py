"""The serving workload at full size: 32 resident trees (n log-uniform in
[1e5, 8e6], d alternating 16 / 64, float32 standard normal promoted to
float64), Poisson arrivals of requests with batch sizes log-uniform in
[1, 65536], k = 10, served with the FP32 screens (matrix cores) and FP64
recheck by the continuous-batching loop.

* Requests overlap on the device: at least two launches are in flight at once.
* Every request's rows equal the same request run alone (kdt_query_device with
  the same flags) byte for byte; for a sample of requests they also equal the
  plain FP64 traversal, exhaustive search over all points of the tree in the
  reference's recipe (d = 16 Folded, d = 64 Streamed; ties by index), and the
  oracle on the smallest trees.
"""

from __future__ import annotations

import numpy as np
import pytest

import recipes
from oracle import OracleTree

pytestmark = pytest.mark.gpu


def tree_points(t, n, d):
    import torch

    from kdtree.workloads import C5

    g = torch.Generator(device="cuda")
    g.manual_seed(C5["seed"] * 1000 + t)
    return torch.randn((n, d), generator=g, device="cuda", dtype=torch.float32).double()


def exhaustive(x, q, k):
    """k nearest of q among the rows of x (device tensors), by the reference
    recipe, ties by index: (reduced distances, indices) on the host."""
    import torch

    qd = torch.from_numpy(q).to(x.device)
    cand_d, cand_i = [], []
    for s in range(0, x.shape[0], 1 << 20):
        red = recipes.reduced(x[s:s + (1 << 20)] - qd)
        kk = min(k, red.shape[0])
        v = torch.topk(red, kk, largest=False).values.max()
        sel = torch.nonzero(red <= v).squeeze(1)
        cand_d.append(red[sel].cpu().numpy())
        cand_i.append((sel + s).cpu().numpy())
    dd, ii = np.concatenate(cand_d), np.concatenate(cand_i)
    order = np.lexsort((ii, dd))[:k]
    return dd[order], ii[order]


def test_served_workload_full_size():
    import torch

    from kdtree import forest as F
    from kdtree.serve import Server, latency_stats, poisson_requests
    from kdtree.workloads import C5, c5_queries, c5_specs

    specs = c5_specs()
    dims = [d for _, d in specs]
    k = C5["k"]
    trees = []
    for t, (n, d) in enumerate(specs):
        x = tree_points(t, n, d)
        trees.append(F.build(x, 32))
        del x
    torch.cuda.synchronize()
    print("built", len(trees), "trees,", sum(n for n, _ in specs), "points", flush=True)
    flags = F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_MATRIX_SCREEN
    try:
        reqs = poisson_requests(240, rate=400.0, tree_dims=dims, k=k, max_batch=C5["max_batch"], seed=C5["seed"],
                                make_queries=c5_queries)
        server = Server(trees, n_streams=4, hbm_budget_bytes=C5["hbm_budget_bytes"], flags=flags)
        server.run(reqs, time_limit=300)
        st = latency_stats(reqs)
        print({"max_inflight": server.max_inflight, "launches": server.launches, **st}, flush=True)
        assert server.max_inflight >= 2, "no two launches were ever in flight together"

        for r in reqs:  # alone, same flags
            q = torch.from_numpy(r.queries).cuda()
            ad, ai = F.query_device(trees[r.tree], q, k, flags=flags | F.KDT_QUERY_ASSUME_FINITE)
            assert np.array_equal(ai.cpu().numpy(), r.out_i), f"request {r.rid}"
            assert ad.cpu().numpy().tobytes() == r.out_d.tobytes(), f"request {r.rid}"
        print("every request equals the request alone", flush=True)

        small = [r for r in reqs if r.queries.shape[0] <= 256][:12]
        for r in small:  # the plain FP64 traversal
            q = torch.from_numpy(r.queries).cuda()
            pd, pi = F.query_device(trees[r.tree], q, k, flags=0)
            assert np.array_equal(pi.cpu().numpy(), r.out_i), f"request {r.rid} plain"
            assert pd.cpu().numpy().tobytes() == r.out_d.tobytes(), f"request {r.rid} plain"
        print(f"{len(small)} requests equal the plain FP64 traversal", flush=True)

        rng = np.random.default_rng(55)
        for r in [reqs[j] for j in rng.choice(len(reqs), 16, replace=False)]:
            n, d = specs[r.tree]
            x = tree_points(r.tree, n, d)
            row = int(rng.integers(0, r.queries.shape[0]))
            red, ind = exhaustive(x, r.queries[row], k)
            del x
            assert np.array_equal(ind, r.out_i[row]), f"request {r.rid} row {row}"
            assert np.sqrt(red).tobytes() == r.out_d[row].tobytes(), f"request {r.rid} row {row}"
        print("exhaustive sample equal", flush=True)

        smallest = sorted(range(len(specs)), key=lambda t: specs[t][0])[:2]
        checked = 0
        for t in smallest:
            host = tree_points(t, *specs[t]).cpu().numpy()
            ot = OracleTree(host, 32)
            for r in [r for r in reqs if r.tree == t][:3]:
                m = min(16, r.queries.shape[0])
                od, oi = ot.query(r.queries[:m], k=k, threads=8)
                assert np.array_equal(oi, r.out_i[:m]) and od.tobytes() == r.out_d[:m].tobytes(), f"request {r.rid}"
                checked += 1
        print(f"oracle equal on {checked} requests", flush=True)
    finally:
        torch.cuda.synchronize()
        for tr in trees:
            tr.free()
