"""Benchmark: exact k-NN queries/s of the plain GPU kernel on one MI355X.

Workload (the plain GPU kernel configuration): a KD-tree over n = 2^20 points,
d = 3, float64 uniform on [0,1)^3, leafsize 32, queried with 2^20 uniform
queries, k = 16, p = 2.  One step = one query pass over the whole batch, with
points, tree and queries resident in HBM.  Results are exact (bit-identical to
the oracle) and a sample is checked each run.

Prints one JSON line (rank 0).  `roofline` uses the algorithmic bytes per
query from the reference traversal counters (leaf rows scanned, nodes visited,
box tests, measured with the oracle on 10^4 queries):
    B_q = 8d + 16k + E(8d + 4) + 24V + 16dX
and the query kernel's average duration from HIP events on its stream.
`cpu_baseline` times the oracle (the C++ restatement of the reference's query,
threaded like its rayon path) on this host's cores.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "kdtree-rs_amd"))

from kdtree._core import load_library  # noqa: E402
from kdtree.workloads import C2, c2_inputs  # noqa: E402

KDT_QUERY_ASSUME_FINITE = 0x2
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5", "screen", "copy"],
                    help="c2 (default, the metric's configuration); c3 precision mode; c4 multi-tree batch; "
                         "c5 serving; screen: the matrix-core screen alone on one 64-d tree; copy: a 1 GiB "
                         "device copy that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--requests", type=int, default=240, help="c5: requests per serving run")
    ap.add_argument("--plain-seconds", type=float, default=40.0, help="c5: time budget of the plain sample")
    return ap.parse_args()


def cpu_model() -> str:
    """The host CPU's model name (first entry of /proc/cpuinfo), or "unknown"."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_profiles():
    """HBM bytes per query launch from the committed rocprofv3 PMC summary, if any."""
    f = ROOT / "profiles" / "knn_c2_pmc.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the workload is a single-GPU "
              "query pass and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c3":
        return run_c3(args)
    if args.workload == "c5":
        return run_c5(args)
    if args.workload == "screen":
        return run_screen(args)
    if args.workload == "copy":
        return run_copy(args)
    if args.workload == "c4":
        return run_c4(args)
    import torch

    lib = load_library()
    dev = torch.device("cuda:0")
    n, nq, d, k, ls = C2["n"], C2["nq"], C2["d"], C2["k"], C2["leafsize"]
    data_h, q_h = c2_inputs()
    data = torch.from_numpy(data_h).to(dev)
    q = torch.from_numpy(q_h).to(dev)
    out_d = torch.empty((nq, k), dtype=torch.float64, device=dev)
    out_i = torch.empty((nq, k), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()

    # Build (timed separately; not part of the metric).
    tree = ctypes.c_void_p()
    t0 = time.perf_counter()
    rc = lib.kdt_build_device(ctypes.c_void_p(data.data_ptr()), n, d, ls, sp, ctypes.byref(tree))
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    if rc:
        raise RuntimeError(lib.kdt_last_error().decode())

    def query(flags):
        r = lib.kdt_query_device(tree, ctypes.c_void_p(q.data_ptr()), nq, d, k, 2.0, float("inf"), 0.0, flags,
                                 ctypes.c_void_p(out_d.data_ptr()), ctypes.c_void_p(out_i.data_ptr()), sp)
        if r:
            raise RuntimeError(lib.kdt_last_error().decode())

    query(0)  # validates the queries (finiteness) once
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        query(KDT_QUERY_ASSUME_FINITE)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        query(KDT_QUERY_ASSUME_FINITE)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    qps = nq * args.steps / wall

    # The drop-in boundary hands over host buffers: the same queries through
    # kdt_query (queries host -> device, finiteness check, query, results
    # device -> host), PCIe included.  Reported beside the metric, never as it.
    hd = np.empty((nq, k), np.float64)
    hi = np.empty((nq, k), np.int64)

    def host_query():
        r = lib.kdt_query(tree, ctypes.c_void_p(q_h.ctypes.data), nq, d, k, 2.0, float("inf"), 0.0, 0,
                          ctypes.c_void_p(hd.ctypes.data), ctypes.c_void_p(hi.ctypes.data))
        if r:
            raise RuntimeError(lib.kdt_last_error().decode())

    host_query()
    t0 = time.perf_counter()
    host_reps = 3
    for _ in range(host_reps):
        host_query()
    host_s = (time.perf_counter() - t0) / host_reps
    host_same = bool(np.array_equal(hi, out_i.cpu().numpy()) and hd.tobytes() == out_d.cpu().numpy().tobytes())

    # Parity sample against the oracle (the checker).
    from oracle import OracleTree

    sample = 2000
    t0 = time.perf_counter()
    otree = OracleTree(data_h, ls)
    cpu_build_ms = (time.perf_counter() - t0) * 1e3  # the oracle's build: one thread
    od, oi, ctr = otree.query(q_h[:10_000], k=k, counters=True)
    gd = out_d[:sample].cpu().numpy()
    gi = out_i[:sample].cpu().numpy()
    parity = bool(np.array_equal(gi, oi[:sample]) and gd.tobytes() == od[:sample].tobytes())

    E, V, X = (ctr / 10_000.0).tolist()
    bq = 8 * d + 16 * k + E * (8 * d + 4) + 24 * V + 16 * d * X
    avg_kernel_s = float(kernel_ms.mean()) / 1e3
    achieved = bq * nq / avg_kernel_s / 1e9
    traffic = traffic_from_profiles()

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        probe = q_h[:20_000]
        t0 = time.perf_counter()
        otree.query(probe, k=k, threads=threads)
        per_q = (time.perf_counter() - t0) / probe.shape[0]
        m = int(min(nq, max(20_000, args.cpu_seconds / max(per_q, 1e-9))))
        t0 = time.perf_counter()
        otree.query(q_h[:m], k=k, threads=threads)
        cpu_qps = m / (time.perf_counter() - t0)
        m1 = int(min(nq, max(2_000, 2.0 / max(per_q * threads, 1e-9))))  # about 2 s on one core
        t0 = time.perf_counter()
        otree.query(q_h[:m1], k=k, threads=1)
        one_core = m1 / (time.perf_counter() - t0)
        cpu = {
            "value": cpu_qps,
            "unit": "queries/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{m} of the {nq} queries, k={k}, on the same tree, {threads} threads "
                      "(oracle = C++ restatement of the reference query, g++ -O3 -ffp-contract=off)",
            "one_core": {"value": one_core, "sample": f"{m1} queries, 1 thread"},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
        }
        try:  # secondary: SciPy's cKDTree (the reference's own tests compare against SciPy)
            import scipy
            from scipy.spatial import cKDTree
        except ImportError:
            scipy = None
        if scipy is not None:
            t0 = time.perf_counter()
            stree = cKDTree(data_h, leafsize=ls)
            sbuild = time.perf_counter() - t0
            t0 = time.perf_counter()
            stree.query(q_h[:m], k=k, workers=threads)
            cpu["scipy_ckdtree"] = {"value": m / (time.perf_counter() - t0), "unit": "queries/s",
                                    "workers": threads, "build_ms": sbuild * 1e3, "version": scipy.__version__,
                                    "sample": f"the same {m} queries and points, leafsize {ls}"}

    line = {
        "metric": "k-NN queries/s (exact indices), plain GPU kernel",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform, seeds 2 / 1002)",
        "config": {"workload": "knn-uniform-n1048576-d3-q1048576-k16-leafsize32-p2", "n_points": n,
                   "n_queries": nq, "ndim": d, "k": k, "leafsize": ls},
        "latency_ms": {"p50": float(np.percentile(kernel_ms, 50)), "p99": float(np.percentile(kernel_ms, 99)),
                       "what": "one full 2^20-query batch"},
        "build_ms": build_ms,
        "build_vs_cpu": {"gpu_ms": build_ms, "cpu_ms": cpu_build_ms, "cpu_threads": 1, "kind": "port",
                         "what": "build of the 2^20-point tree: oracle on one host thread (the build parity "
                                 "tests check that both trees are identical); GPU time includes its allocations"},
        "parity_sample_bit_exact": parity,
        "host_buffers_pcie_inclusive": {"queries_per_s": nq / host_s, "ms_per_call": host_s * 1e3,
                                        "same_as_device_path": host_same,
                                        "what": "kdt_query on pageable host arrays: 24 MiB of queries in, "
                                                "256 MiB of results out, finiteness check included"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_query": bq, "E": E, "V": V, "X": X,
                     "kernel_ms_avg": float(kernel_ms.mean())},
        "distance_evals_per_s": E * nq / avg_kernel_s,
        "cpu_baseline": cpu,
    }
    lib.kdt_free(tree)
    print(json.dumps(line), flush=True)


def _time_steps(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def run_c3(args):
    """Precision mode: n = 2^22, d = 8, 64-component Gaussian mixture, 10^6
    queries, k = 32; all-FP64 leaf scan versus FP32 screen + FP64 recheck."""
    import torch

    from kdtree import forest as F
    from kdtree.workloads import C3, c3_inputs
    from oracle import OracleTree

    data_h, q_h = c3_inputs()
    k, ls = C3["k"], C3["leafsize"]
    data = torch.from_numpy(data_h).cuda()
    q = torch.from_numpy(q_h).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = F.build(data, ls)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    nq = q.shape[0]
    od64, oi64 = F.query_device(tree, q, k, flags=0)
    od32, oi32 = F.query_device(tree, q, k, flags=F.KDT_QUERY_FP32_SCREEN)
    s64 = _time_steps(lambda: F.query_device(tree, q, k, od64, oi64), args.steps, args.warmup)
    s32 = _time_steps(lambda: F.query_device(tree, q, k, od32, oi32,
                                             flags=F.KDT_QUERY_ASSUME_FINITE | F.KDT_QUERY_FP32_SCREEN),
                      args.steps, args.warmup)
    same_idx = bool(torch.equal(oi64, oi32))
    d64, d32 = od64.cpu().numpy(), od32.cpu().numpy()
    rel = float(np.max(np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-300)))
    sample = 2000
    t0 = time.perf_counter()
    otree = OracleTree(data_h, ls)
    cpu_build_ms = (time.perf_counter() - t0) * 1e3  # the oracle's build: one thread
    od, oi = otree.query(q_h[:sample], k=k)
    parity = bool(np.array_equal(oi, oi64[:sample].cpu().numpy()) and od.tobytes() == d64[:sample].tobytes())
    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        m = 20_000
        t0 = time.perf_counter()
        otree.query(q_h[:m], k=k, threads=threads)
        cpu = {"value": m / (time.perf_counter() - t0), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"{m} of the {nq} queries, k={k}, {threads} threads (oracle)"}
    F_tree = tree
    out = {
        "metric": "k-NN queries/s (exact indices), precision mode",
        "config": {"workload": "knn-gmm64-n4194304-d8-q1000000-k32-leafsize32-p2"},
        "fp64": {"value": nq / s64, "ms_per_step": s64 * 1e3},
        "fp32_screen": {"value": nq / s32, "ms_per_step": s32 * 1e3},
        "indices_identical": same_idx, "max_rel_distance_error_fp32_vs_fp64": rel,
        "parity_sample_bit_exact_vs_oracle": parity, "build_ms": build_ms, "cpu_baseline": cpu,
        "vs_cpu_fp32_screen": (nq / s32) / cpu["value"] if cpu else None,
    }
    F_tree.free()
    print(json.dumps(out), flush=True)


def run_c4(args):
    """256 trees (n log-uniform in [1e3,1e6], d in {2,3,8,16}, float32 uniform
    promoted to float64), per-tree query counts log-uniform in [1, 1e5], k in
    {1, 8, 32}; leafsize autotuned over {8,16,32,64}, lanes per query over the
    widths that hold k, and the matrix-core screen where it applies.  Shared
    persistent launches (items of 8192+ queries on their own) versus one launch
    sequence per tree versus the per-tree launches captured as a HIP graph (a
    pure enqueue: no leaf ordering and no matrix-core screen)."""
    import torch

    from kdtree import forest as F
    from kdtree.workloads import c4_specs, c4_tree_inputs
    from oracle import OracleTree

    specs = c4_specs()
    items, host = [], []
    t0 = time.perf_counter()
    for t, (n, d, q, k) in enumerate(specs):
        data, queries = c4_tree_inputs(t, n, d, q)
        tree = F.autotune_build(data, queries[:4096], k, reps=2)
        items.append((tree, torch.from_numpy(queries).cuda(), k))
        if t < 8:
            host.append((data, queries, k, tree.leafsize))
    tune_s = time.perf_counter() - t0
    batch = F.Batch(items)
    nq = batch.n_queries
    res, outs = {}, {}
    for mode in ("per_tree", "graph", "persistent"):
        res[mode] = nq / _time_steps(lambda: batch.run(mode), args.steps, args.warmup)
        outs[mode] = [(od.clone(), oi.clone()) for _, _, _, od, oi in batch.items]
    modes_equal = all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
                      for m in ("per_tree", "graph") for a, b in zip(outs[m], outs["persistent"]))
    batch.run("persistent")
    torch.cuda.synchronize()
    parity = True
    for j, (data, queries, k, ls) in enumerate(host):
        od, oi = OracleTree(data, ls).query(queries, k=k)
        gd, gi = batch.results(j)
        parity &= bool(np.array_equal(gi.cpu().numpy(), oi) and gd.cpu().numpy().tobytes() == od.tobytes())
    leafs, groups = {}, {}
    for tree, _, _ in items:
        leafs[tree.leafsize] = leafs.get(tree.leafsize, 0) + 1
        groups[tree.group] = groups.get(tree.group, 0) + 1
    out = {
        "metric": "k-NN queries/s over a 256-tree batch",
        "config": {"workload": "knn-256-trees-n1e3..1e6-d{2,3,8,16}-k{1,8,32}", "n_queries": nq,
                   "n_points": int(sum(s[0] for s in specs))},
        "queries_per_s": res,
        "persistent_vs_per_tree": res["persistent"] / res["per_tree"],
        "autotuned_leafsize_counts": leafs, "autotuned_lanes_per_query_counts": groups, "autotune_s": tune_s,
        "parity_first_8_trees_bit_exact": parity,
        "all_items_bitwise_equal_across_modes": bool(modes_equal),
        "matrix_screen_trees": sum(1 for tree, _, _ in items if tree.flags),
    }
    for tree, _, _ in items:
        tree.free()
    print(json.dumps(out), flush=True)


def run_copy(args):
    """3 copies of 1 GiB (device to device): 1 GiB read + 1 GiB written per
    dispatch, the known traffic that scales the PMC byte counters."""
    import torch

    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_dispatch": 1 << 30, "bytes_written_per_dispatch": 1 << 30,
                      "dispatches": 3}), flush=True)


def run_screen(args):
    """The matrix-core screen on its own: n = 2^20 float32 standard normal
    rows of d = 64 (as float64), 65536 queries, k = 10.  Roofline: the bf16
    matrix-core peak, with the split contraction's 3 x 2 d bf16 FLOPs per
    (query, point) pair (qh.ph + ql.ph + qh.pl) as the algorithmic work; the
    line also gives the 2 d FLOPs per pair of the plain contraction."""
    import torch

    from kdtree import forest as F
    from oracle import OracleTree

    n, d, nq, k = 1 << 20, 64, 1 << 16, 10
    g = torch.Generator(device="cuda")
    g.manual_seed(64)
    data = torch.randn((n, d), generator=g, device="cuda", dtype=torch.float32).double()
    q = torch.randn((nq, d), generator=g, device="cuda", dtype=torch.float32).double()
    tree = F.build(data, 32)
    flags = F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_ASSUME_FINITE
    od, oi = F.query_device(tree, q, k, flags=flags)
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    for _ in range(args.warmup):
        F.query_device(tree, q, k, od, oi, flags=flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        F.query_device(tree, q, k, od, oi, flags=flags)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = np.array([s_.elapsed_time(e_) for s_, e_ in zip(starts, ends)])
    flop = 3 * 2.0 * d * n * nq
    achieved = flop / (ms.mean() / 1e3) / 1e12
    peak = 2500.0  # dense bf16 matrix-core peak, TFLOP/s
    m = 64
    otree = OracleTree(data.cpu().numpy(), 32)
    pd, pi = otree.query(q[:m].cpu().numpy(), k=k, threads=8)
    parity = bool(np.array_equal(pi, oi[:m].cpu().numpy()) and pd.tobytes() == od[:m].cpu().numpy().tobytes())
    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        qs = q[:256].cpu().numpy()
        t1 = time.perf_counter()
        otree.query(qs, k=k, threads=threads)
        cpu = {"value": qs.shape[0] / (time.perf_counter() - t1), "unit": "queries/s", "cores": threads,
               "kind": "port", "sample": f"256 of the {nq} queries, {threads} threads (oracle)"}
    tree.free()
    print(json.dumps({
        "metric": "k-NN queries/s (exact indices), matrix-core screen, 64-d",
        "value": nq * args.steps / wall, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "dtype": "split bf16 (matrix cores, f32 accumulate) + f64 recheck", "data": "synthetic (torch.randn float32, seed 64)",
        "config": {"workload": "knn-normal-n1048576-d64-q65536-k10-leafsize32-p2"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "what": "3 x 2 d bf16 FLOPs per (query, point) pair over the whole screened query "
                             "call (matrix-core screen, candidate selection, FP64 recheck, certificate)",
                     "plain_contraction_tflops": achieved / 3},
        "parity_sample_bit_exact": parity, "cpu_baseline": cpu,
    }), flush=True)


def run_c5(args):
    """Serving: 32 resident trees (n log-uniform in [1e5, 8e6], d alternating
    16 / 64, float32 standard normal), Poisson requests on uniform trees with
    batch sizes log-uniform in [1, 65536], k = 10, 200 GB HBM admission budget.

    plain   each request alone, one after another, through kdt_query (host
            buffers in and out, the FP64 traversal kernel): what a caller of
            the reference API gets from the plain GPU kernel; and the same
            requests with their queries resident in HBM, one kdt_query_device
            call per tree (the plain kernel alone);
    served  the continuous-batching loop (kdtree.serve.Server): requests are
            admitted while others run, coalesced into multi-tree launches on
            several streams, FP32 screens (the matrix-core screen for d = 64,
            and for d = 16 when it measures faster) with FP64 recheck.
    Both give identical results; a sample is checked against the oracle and
    against the plain path.  Latency = arrival to host-visible result."""
    import torch

    from kdtree import KDTree  # noqa: F401  (loads the library with torch's runtime first)
    from kdtree import forest as F
    from kdtree.serve import Request, Server, latency_stats, poisson_requests
    from kdtree.workloads import C5, c5_queries, c5_specs
    from oracle import OracleTree

    lib = load_library()
    specs = c5_specs()
    dims = [d for _, d in specs]
    k = C5["k"]
    smallest = sorted(range(len(specs)), key=lambda t: specs[t][0])[:3]
    trees, host_data = [], {}
    t0 = time.perf_counter()
    for t, (n, d) in enumerate(specs):
        g = torch.Generator(device="cuda")
        g.manual_seed(C5["seed"] * 1000 + t)
        x = torch.randn((n, d), generator=g, device="cuda", dtype=torch.float32).double()
        trees.append(F.build(x, 32))
        if t in smallest:
            host_data[t] = x.cpu().numpy()
        del x
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0

    # d = 16: matrix-core screen or traversal, whichever is faster on a sample.
    t16 = [t for t in range(len(specs)) if dims[t] == 16]
    t16 = sorted(t16, key=lambda t: specs[t][0])[len(t16) // 2]
    qs = torch.from_numpy(c5_queries(10**6, 4096, 16)).cuda()
    tune = {}
    for name, fl in (("traversal_fp32_screen", F.KDT_QUERY_FP32_SCREEN),
                     ("matrix_screen", F.KDT_QUERY_FP32_SCREEN | 0x8)):
        F.query_device(trees[t16], qs, k, flags=fl | F.KDT_QUERY_ASSUME_FINITE)
        tune[name] = _time_steps(lambda: F.query_device(trees[t16], qs, k, flags=fl | F.KDT_QUERY_ASSUME_FINITE),
                                 3, 1) * 1e3
    serve_flags = F.KDT_QUERY_FP32_SCREEN | (0x8 if tune["matrix_screen"] < tune["traversal_fp32_screen"] else 0)

    base = poisson_requests(args.requests, rate=1.0, tree_dims=dims, k=k, max_batch=C5["max_batch"],
                            seed=C5["seed"], make_queries=c5_queries)

    def fresh(rate, subset=None):
        src = base if subset is None else subset
        return [Request(r.rid, r.arrival / rate, r.tree, r.queries, r.k) for r in src]

    # Plain: one request at a time, in arrival order, within the time budget;
    # each request cut to at most 2048 queries (the served comparison on the
    # same sample runs the same cut requests).
    plain, plain_t, plain_q = [], 0.0, 0
    for r0 in base:
        r = Request(r0.rid, r0.arrival, r0.tree, np.ascontiguousarray(r0.queries[:2048]), r0.k)
        out_d = np.empty((r.queries.shape[0], k))
        out_i = np.empty((r.queries.shape[0], k), np.int64)
        s0 = time.perf_counter()
        rc = lib.kdt_query(trees[r.tree].handle, ctypes.c_void_p(r.queries.ctypes.data), r.queries.shape[0],
                           r.queries.shape[1], k, 2.0, float("inf"), 0.0, 0, ctypes.c_void_p(out_d.ctypes.data),
                           ctypes.c_void_p(out_i.ctypes.data))
        dt = time.perf_counter() - s0
        assert rc == 0, lib.kdt_last_error().decode()
        plain.append((r, dt, out_d, out_i))
        plain_t += dt
        plain_q += r.queries.shape[0]
        if plain_t > args.plain_seconds and len(plain) >= 4:
            break
    plain_lat = np.array([dt for _, dt, _, _ in plain]) * 1e3
    sample = [r for r, _, _, _ in plain]

    # Plain kernel, device-resident: the same sample, each tree's requests
    # concatenated into one kdt_query_device call (the FP64 traversal, no
    # screen, queries already in HBM): the plain GPU kernel without the
    # per-request host copies and calls.
    by_tree = {}
    for j, r in enumerate(sample):
        by_tree.setdefault(r.tree, []).append(j)
    dev = {t: torch.from_numpy(np.concatenate([sample[j].queries for j in js])).cuda() for t, js in by_tree.items()}
    for t in dev:
        F.query_device(trees[t], dev[t][:1], k)
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    pk_out = {t: F.query_device(trees[t], dev[t], k) for t in dev}
    torch.cuda.synchronize()
    pk_t = time.perf_counter() - s0
    pk_exact = all(
        np.array_equal(pk_out[t][1].cpu().numpy(), np.concatenate([plain[j][3] for j in js]))
        and pk_out[t][0].cpu().numpy().tobytes() == np.concatenate([plain[j][2] for j in js]).tobytes()
        for t, js in by_tree.items())
    del dev, pk_out

    def serve(reqs):
        server = Server(trees, n_streams=4, hbm_budget_bytes=C5["hbm_budget_bytes"],
                        max_queries_per_launch=1 << 18, flags=serve_flags)
        server.run(reqs, time_limit=600)
        st = latency_stats(reqs)
        st.update(launches=server.launches, max_inflight=server.max_inflight)
        return st

    # Served on the plain sample (same work), then the full stream at saturation
    # (every request already waiting) and at fractions of that capacity.
    same = fresh(1e12, sample)
    served_sample = serve(same)
    exact_vs_plain = all(np.array_equal(s.out_i, pi) and s.out_d.tobytes() == pd.tobytes()
                         for s, (_, _, pd, pi) in zip(same, plain))
    sat = serve(fresh(1e12))
    cap_rps = sat["requests"] / sat["span_s"]
    loads = {}
    for frac in (0.5, 0.8):
        loads[f"{frac:.1f}x_capacity"] = serve(fresh(cap_rps * frac))

    parity = True
    checked = 0
    full = fresh(1e12)
    serve(full)
    for r in full:
        if r.tree in host_data and checked < 4:
            m = min(16, r.queries.shape[0])
            od, oi = OracleTree(host_data[r.tree], 32).query(r.queries[:m], k=k, threads=8)
            parity &= bool(np.array_equal(oi, r.out_i[:m]) and od.tobytes() == r.out_d[:m].tobytes())
            checked += 1
    out = {
        "metric": "k-NN queries/s, served (continuous batching + FP32/matrix-core screen) vs plain GPU kernel, "
                  "exact indices",
        "value": sat["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32 screen + f64 recheck", "data": "synthetic (torch.randn float32, seed 5000+t; SplitMix64 "
                                                     "Box-Muller queries)",
        "config": {"workload": "serve-32-trees-n1e5..8e6-d{16,64}-k10-batch1..65536-budget200GB",
                   "n_points": int(sum(n for n, _ in specs)), "requests": args.requests},
        "plain": {"queries_per_s": plain_q / plain_t, "requests": len(plain), "queries": plain_q,
                  "note": "first requests in arrival order, each cut to <= 2048 queries",
                  "p50_ms": float(np.percentile(plain_lat, 50)), "p99_ms": float(np.percentile(plain_lat, 99))},
        "plain_kernel_device_resident": {"queries_per_s": plain_q / pk_t, "seconds": pk_t,
                                         "note": "same sample, one kdt_query_device call per tree, FP64 traversal",
                                         "exact_vs_plain_bitwise": bool(pk_exact)},
        "served_same_requests": served_sample,
        "served_vs_plain_same_requests": served_sample["queries_per_s"] / (plain_q / plain_t),
        "served_vs_plain_kernel_same_requests": served_sample["queries_per_s"] / (plain_q / pk_t),
        "served_saturation": sat, "served_at_load": loads,
        "d16_choice_ms_per_4096_queries": tune, "d16_matrix_screen": bool(serve_flags & 0x8),
        "exact_vs_plain_bitwise": exact_vs_plain, "parity_vs_oracle_sample": parity, "oracle_checked_requests": checked,
        "build_s": build_s,
    }
    for tr in trees:
        tr.free()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
