"""The batching workload at full size: 256 independent trees (n log-uniform in
[1e3, 1e6], d from {2, 3, 8, 16}, float32 uniform points promoted to float64),
per-tree query counts log-uniform in [1, 1e5], k from {1, 8, 32}, leafsize
autotuned per tree over {8, 16, 32, 64}.

Checked through properties that hold at any size, as the reference's tests
check every result (tests/test_kdtree.py:55-89) and its execution modes give
identical trees and answers (src/build.rs:345-359):

* every tree: the exported indices are a permutation of 0..n-1 and every tree
  row is the input row it names, bit for bit;
* every result row of every item: indices distinct and in range, sorted by
  (distance, index), and every distance equal to the reference's FP64 recipe
  for its width recomputed here (for d = 8 the chunk rule of the leaf the point
  sits in, which depends on the autotuned leafsize);
* on a sample of queries per tree, exhaustive search over all points gives the
  same k nearest, ties by index; on the first trees the oracle gives the same
  bytes;
* the shared persistent launches, per-tree launches, the batch plan and its
  HIP graph replays give bitwise identical outputs.
"""

from __future__ import annotations

import numpy as np
import pytest

import recipes
from oracle import OracleTree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    import torch

    from kdtree import forest as F
    from kdtree.workloads import c4_specs, c4_tree_inputs

    specs = c4_specs()
    items, chosen, host = [], [], []
    for t, (n, d, q, k) in enumerate(specs):
        data, queries = c4_tree_inputs(t, n, d, q)
        host.append((data, queries))
        tree = F.autotune_build(data, queries[:1024], k, reps=1)
        items.append((tree, torch.from_numpy(queries).cuda(), k))
        chosen.append((tree.leafsize, tree.group, tree.flags))
        if t % 32 == 31:
            print(f"built and tuned {t + 1} trees", flush=True)
    batch = F.Batch(items)
    print("leafsize counts", {ls: sum(1 for c in chosen if c[0] == ls) for ls in (8, 16, 32, 64)},
          "matrix-screen trees", sum(1 for c in chosen if c[2]), "queries", batch.n_queries, flush=True)
    yield specs, batch, host
    batch.close()
    for tree, _, _ in items:
        tree.free()


def snapshot(batch):
    return [(od.cpu().numpy().copy(), oi.cpu().numpy().copy()) for _, _, _, od, oi in batch.items]


def test_modes_bitwise_equal(c4):
    import torch

    _, batch, _ = c4
    outs = {}
    for mode in ("persistent", "per_tree", "plan", "graph", "graph", "graph"):
        for _, _, _, od, oi in batch.items:
            od.fill_(-1.0)
            oi.fill_(-1)
        torch.cuda.synchronize()
        batch.run(mode)
        torch.cuda.synchronize()
        snap = snapshot(batch)
        if "persistent" in outs:
            for j, ((a, b), (c, d)) in enumerate(zip(snap, outs["persistent"])):
                assert a.tobytes() == c.tobytes() and np.array_equal(b, d), f"{mode} item {j}"
        outs.setdefault(mode, snap)
        print(f"mode {mode} ok", flush=True)


def test_every_row_against_the_recipe_and_exhaustive_sample(c4):
    import torch

    from kdtree import forest as F

    specs, batch, host = c4
    batch.run("persistent")
    torch.cuda.synchronize()
    rng = np.random.default_rng(44)
    for t, ((n, d, q, k), (tree, _, _, od, oi)) in enumerate(zip(specs, batch.items)):
        data, queries = host[t]
        rows, idx, nodes = F.export(tree)
        perm = idx.astype(np.int64)
        assert np.array_equal(np.sort(perm), np.arange(n)), f"tree {t}"
        assert rows.tobytes() == data[perm].tobytes(), f"tree {t}"
        dist, ind = od.cpu().numpy(), oi.cpu().numpy()
        assert ind.min() >= 0 and ind.max() < n, f"tree {t}"
        dd, di = dist[:, 1:] - dist[:, :-1], ind[:, 1:] - ind[:, :-1]
        assert np.all((dd > 0) | ((dd == 0) & (di > 0))), f"tree {t}: rows not sorted by (distance, index)"
        full = None
        if d == 8:
            pos = np.empty(n, np.int64)
            pos[perm] = np.arange(n)
            full_pos = recipes.leaf_full_mask(nodes, n, 8)
        for s in range(0, q, 8192):
            e = min(q, s + 8192)
            if d == 8:
                full = full_pos[pos[ind[s:e]]]
            ref = np.sqrt(recipes.reduced(data[ind[s:e]] - queries[s:e, None, :], full))
            assert ref.tobytes() == dist[s:e].tobytes(), f"tree {t} rows {s}..{e}"
        # exhaustive search on a sample of queries
        if d == 8:
            full_all = full_pos[pos]
        for r in rng.choice(q, min(q, 3), replace=False):
            all_d = np.sqrt(recipes.reduced(data - queries[r], full_all if d == 8 else None))
            order = np.lexsort((np.arange(n), all_d))[:k]
            assert np.array_equal(ind[r], order), f"tree {t} query {r}"
            assert all_d[order].tobytes() == dist[r].tobytes(), f"tree {t} query {r}"
        if t < 16:  # the oracle, with the autotuned leafsize
            m = min(q, 64)
            odist, oind = OracleTree(data, tree.leafsize).query(queries[:m], k=k, threads=8)
            assert np.array_equal(oind, ind[:m]) and odist.tobytes() == dist[:m].tobytes(), f"tree {t} oracle"
        if t % 32 == 31:
            print(f"checked {t + 1} trees", flush=True)
