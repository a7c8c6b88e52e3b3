"""Continuous-batching k-NN server over resident trees (one GPU).

No reference counterpart: the reference is a library call per query matrix.
Here trees stay resident in HBM and requests (tree, query rows, k) arrive over
time.  The dispatcher loop:

  * admits every request whose arrival time has passed into the pending queue;
  * retires launches whose completion event has fired (results are then in
    pinned host memory) and frees their share of the HBM admission budget;
  * while a stream is free and the budget allows, takes pending requests in
    arrival order up to the per-launch query cap and dispatches them as ONE
    multi-tree launch (kdt_query_batch_device) on that stream: the requests
    on one tree (with one k) are concatenated into one item, queries go
    host->device, the launch runs, results go device->host, then an event;
  * the first stream is a lane for small requests (at most `small_max`
    queries): when it is free it takes every pending small request, so that a
    small request never waits behind launches of large ones; the other
    streams take every request in arrival order;
  * a request of more than `slice_max` queries is admitted as slices of at
    most `slice_max` rows, and a launch takes at most one slice of each
    request: a large request then spreads over several launches (and
    streams) instead of holding one stream for its whole length, so the
    longest launch, and with it the wait of a request that arrives behind
    it, stays bounded.  The request completes with its last slice.

Requests that arrive while earlier launches run are admitted at the next
dispatch and run concurrently on another stream, so work from different
requests overlaps on the device.  Every request's rows are computed by the
same per-query descent whatever else is in the launch, so a result is the
same alone, sliced and in any mix.  Latency = host-visible completion -
arrival.
"""

from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import forest as F


@dataclass
class Request:
    rid: int
    arrival: float  # seconds from the start of the run
    tree: int
    queries: np.ndarray  # (m, d) float64
    k: int
    done: float = -1.0
    out_d: np.ndarray | None = None
    out_i: np.ndarray | None = None


@dataclass
class Slice:
    """Rows [lo, hi) of a request: the unit a launch takes."""

    req: Request
    lo: int
    hi: int

    @property
    def queries(self) -> np.ndarray:
        return self.req.queries[self.lo:self.hi]

    @property
    def rid(self) -> int:
        return self.req.rid


def split_request(r: Request, slice_max: int) -> list:
    """The slices of `r`, at most `slice_max` rows each (one slice if it fits)."""
    m = r.queries.shape[0]
    return [Slice(r, lo, min(m, lo + slice_max)) for lo in range(0, m, slice_max)]


@dataclass
class _Launch:
    stream: object
    event: object
    groups: list  # (slices, pinned host distances, pinned host indices), one per item
    dev: list = field(default_factory=list)
    nbytes: int = 0


class Server:
    def __init__(self, trees, n_streams: int = 4, hbm_budget_bytes: float = 200e9,
                 max_queries_per_launch: int = 1 << 18, p: float = 2.0, flags: int = 0,
                 small_max: int = 4096, slice_max: int = 8192):
        import torch

        self.torch = torch
        self.trees = trees
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]  # default priority
        self.budget = hbm_budget_bytes
        self.cap = max_queries_per_launch
        self.p = p
        self.flags = flags
        self.small_max = small_max
        self.slice_max = slice_max
        self.lib = F._lib()
        self.max_inflight = 0
        self.launches = 0

    def _bytes(self, sl) -> int:
        """Device bytes a slice holds while in flight: queries, outputs and
        an upper bound of the matrix-core screen's scratch (two 16-entry lists
        of 8-byte entries per query and point slice, with the launch's slice
        rule: enough slices for 4096 query-tile waves, at least k / 4 for
        k > 16, at most 512; then 64 candidates, a count and a flag per query)."""
        m, d = sl.queries.shape
        k = sl.req.k
        qtiles = -(-m // 32)
        slices = max(-(-4096 // qtiles), -(-k // 4) if k > 16 else 1)
        slices = min(slices, 512)
        return m * d * 8 + m * k * 16 + m * (256 * slices + 261)

    def _dispatch(self, slices, stream) -> _Launch:
        torch = self.torch
        by_item: dict = {}
        for sl in slices:  # slices on one tree with one k share an item
            by_item.setdefault((sl.req.tree, sl.req.k), []).append(sl)
        items = (F.BatchItem * len(by_item))()
        dev, groups = [], []
        with torch.cuda.stream(stream):
            for j, ((tree_id, k), rs) in enumerate(by_item.items()):
                q = np.ascontiguousarray(rs[0].queries) if len(rs) == 1 else np.concatenate([x.queries for x in rs])
                m = q.shape[0]
                qh = torch.from_numpy(q).pin_memory()
                qd = qh.to("cuda", non_blocking=True)
                od = torch.empty((m, k), dtype=torch.float64, device="cuda")
                oi = torch.empty((m, k), dtype=torch.int64, device="cuda")
                t = self.trees[tree_id]
                items[j] = F.BatchItem(t.handle.value, qd.data_ptr(), m, t.d, k, od.data_ptr(), oi.data_ptr())
                dev.append((qh, qd, od, oi))
            rc = self.lib.kdt_query_batch_device(items, len(by_item), self.p, float("inf"), 0.0, self.flags,
                                                 ctypes.c_void_p(stream.cuda_stream))
            if rc:
                raise RuntimeError(self.lib.kdt_last_error().decode())
            for rs, (_, _, od, oi) in zip(by_item.values(), dev):
                hd = torch.empty(od.shape, dtype=torch.float64, pin_memory=True)
                hi = torch.empty(oi.shape, dtype=torch.int64, pin_memory=True)
                hd.copy_(od, non_blocking=True)
                hi.copy_(oi, non_blocking=True)
                groups.append((rs, hd, hi))
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Launch(stream, ev, groups, dev, sum(self._bytes(x) for x in slices))

    def run(self, requests, time_limit: float = 600.0):
        """Serve `requests` (sorted by arrival); returns them with results and completion times."""
        pending: deque = deque()
        upcoming = deque(sorted(requests, key=lambda r: r.arrival))
        inflight: list[_Launch] = []
        free = list(self.streams)
        left: dict = {}  # request id -> slices not yet retired
        used = 0
        t0 = time.perf_counter()
        while upcoming or pending or inflight:
            now = time.perf_counter() - t0
            if now > time_limit:
                raise TimeoutError("serving run exceeded its time limit")
            while upcoming and upcoming[0].arrival <= now:
                r = upcoming.popleft()
                m = r.queries.shape[0]
                r.out_d = np.empty((m, r.k), np.float64)
                r.out_i = np.empty((m, r.k), np.int64)
                parts = split_request(r, self.slice_max)
                left[id(r)] = len(parts)
                pending.extend(parts)
            for L in list(inflight):
                if L.event.query():
                    tdone = time.perf_counter() - t0
                    for sls, hd, hi in L.groups:
                        d_all, i_all = hd.numpy(), hi.numpy()
                        off = 0
                        for sl in sls:
                            m = sl.hi - sl.lo
                            sl.req.out_d[sl.lo:sl.hi] = d_all[off:off + m]
                            sl.req.out_i[sl.lo:sl.hi] = i_all[off:off + m]
                            off += m
                            left[id(sl.req)] -= 1
                            if left[id(sl.req)] == 0:
                                sl.req.done = tdone
                    inflight.remove(L)
                    free.append(L.stream)
                    used -= L.nbytes
            dispatched = False
            while pending and free:
                take, stream, pending = plan_launch(pending, free, self.streams[0], self.small_max, self.cap,
                                                    self.budget - used, bool(inflight), self._bytes)
                if not take:
                    break
                free.remove(stream)
                L = self._dispatch(take, stream)
                inflight.append(L)
                used += L.nbytes
                self.launches += 1
                self.max_inflight = max(self.max_inflight, len(inflight))
                dispatched = True
            if not dispatched:
                time.sleep(0.0002)
        return requests


def plan_launch(pending, free, lane, small_max, cap, budget_left, busy, nbytes):
    """The next launch: (units, stream, units still pending).

    `pending` holds work units (requests or request slices, each with
    `queries` and `rid`) in arrival order, `free` the idle streams, `lane` the
    stream kept for small units (at most `small_max` queries).  When the lane
    is idle and a small unit waits, the launch takes pending small units only,
    on the lane; otherwise it takes units in arrival order on another idle
    stream.  A launch takes at most one unit of each request (the other slices
    of a request wait for a later launch).  It stops at `cap` queries (the
    first unit always fits) and at `budget_left` bytes of `nbytes(unit)`,
    which a launch may exceed only when nothing else runs (`busy` false).
    Returns no units when nothing can start."""
    small_lane = lane in free and any(r.queries.shape[0] <= small_max for r in pending)
    others = [x for x in free if x is not lane]
    if not small_lane and not others:
        return [], None, pending
    pending = deque(pending)
    take, nq, nb, keep, rids = [], 0, 0, deque(), set()
    while pending and (not take or nq + pending[0].queries.shape[0] <= cap):
        r = pending[0]
        if (small_lane and r.queries.shape[0] > small_max) or r.rid in rids:
            keep.append(pending.popleft())  # for the other streams, or a later launch
            continue
        b = nbytes(r)
        if nb + b > budget_left and (take or busy):
            break
        take.append(pending.popleft())
        rids.add(r.rid)
        nq += r.queries.shape[0]
        nb += b
    keep.extend(pending)
    return take, (lane if small_lane else others[0]) if take else None, keep


def latency_stats(requests):
    lat = np.array([r.done - r.arrival for r in requests]) * 1e3
    nq = sum(r.queries.shape[0] for r in requests)
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "queries_per_s": nq / span, "requests": len(requests),
            "queries": int(nq), "span_s": span}


def poisson_requests(n_requests: int, rate: float, tree_dims, k: int, max_batch: int, seed: int,
                     make_queries):
    """Arrivals at `rate` per second; tree uniform; batch size log-uniform in [1, max_batch]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    out = []
    for i in range(n_requests):
        tree = int(rng.integers(0, len(tree_dims)))
        m = max(1, int(round(10 ** (np.log10(max_batch) * rng.uniform()))))
        out.append(Request(i, float(t[i]), tree, make_queries(i, m, tree_dims[tree]), k))
    return out
