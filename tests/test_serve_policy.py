"""The serving loop's launch planning (kdtree.serve.plan_launch), on the CPU:
which requests the next launch takes and on which stream."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from kdtree.serve import plan_launch

LANE, S1, S2 = "lane", "s1", "s2"


def reqs(*sizes):
    return [SimpleNamespace(rid=i, queries=np.empty((m, 1))) for i, m in enumerate(sizes)]


def plan(pending, free, small_max=100, cap=1000, budget_left=10**9, busy=False, nbytes=lambda r: 1):
    take, stream, rest = plan_launch(pending, free, LANE, small_max, cap, budget_left, busy, nbytes)
    return [r.rid for r in take], stream, [r.rid for r in rest]


def test_idle_lane_takes_only_small_requests_in_arrival_order():
    assert plan(reqs(50, 500, 10, 700, 100), [LANE, S1]) == ([0, 2, 4], LANE, [1, 3])


def test_busy_lane_leaves_arrival_order_to_the_other_streams():
    assert plan(reqs(50, 500, 10), [S1, S2]) == ([0, 1, 2], S1, [])


def test_idle_lane_with_only_large_requests_waiting_defers_to_other_streams():
    assert plan(reqs(500, 300), [LANE, S2]) == ([0, 1], S2, [])
    assert plan(reqs(500, 300), [LANE]) == ([], None, [0, 1])


def test_query_cap_and_an_oversized_first_request():
    assert plan(reqs(400, 400, 400), [S1], cap=1000) == ([0, 1], S1, [2])
    assert plan(reqs(5000, 10), [S1], cap=1000) == ([0], S1, [1])


def test_hbm_budget():
    size = lambda r: r.queries.shape[0]  # noqa: E731
    # over budget: nothing starts while other launches hold memory ...
    assert plan(reqs(500), [S1], budget_left=100, busy=True, nbytes=size) == ([], None, [0])
    # ... but one request may start when nothing else runs
    assert plan(reqs(500, 20), [S1], budget_left=100, busy=False, nbytes=size) == ([0], S1, [1])
    assert plan(reqs(30, 40, 50), [S1], budget_left=100, busy=True, nbytes=size) == ([0, 1], S1, [2])


def test_one_unit_per_request_per_launch():
    # Units 0 and 2 are slices of one request (rid 7): the launch takes the
    # first, the second waits for a later launch, in its place.
    units = [SimpleNamespace(rid=7, queries=np.empty((100, 1))), SimpleNamespace(rid=8, queries=np.empty((50, 1))),
             SimpleNamespace(rid=7, queries=np.empty((30, 1))), SimpleNamespace(rid=9, queries=np.empty((20, 1)))]
    take, stream, rest = plan_launch(units, [S1], LANE, 10, 1000, 10**9, False, lambda r: 1)
    assert [(u.rid, u.queries.shape[0]) for u in take] == [(7, 100), (8, 50), (9, 20)]
    assert stream == S1 and [(u.rid, u.queries.shape[0]) for u in rest] == [(7, 30)]


def test_split_request_covers_rows_once():
    from kdtree.serve import Request, split_request

    r = Request(3, 0.0, 0, np.arange(20_000, dtype=np.float64).reshape(-1, 1), 10)
    parts = split_request(r, 8192)
    assert [(p.lo, p.hi) for p in parts] == [(0, 8192), (8192, 16384), (16384, 20_000)]
    assert np.array_equal(np.concatenate([p.queries for p in parts]), r.queries)
    assert all(p.rid == 3 for p in parts)
    small = Request(4, 0.0, 0, np.zeros((5, 2)), 10)
    assert [(p.lo, p.hi) for p in split_request(small, 8192)] == [(0, 5)]
