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
    return ap.parse_args()


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

    # Parity sample against the oracle (the checker).
    from oracle import OracleTree

    sample = 2000
    otree = OracleTree(data_h, ls)
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
        cpu = {
            "value": cpu_qps,
            "unit": "queries/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{m} of the {nq} queries, k={k}, on the same tree, {threads} threads "
                      "(oracle = C++ restatement of the reference query, g++ -O3 -ffp-contract=off)",
        }

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
        "parity_sample_bit_exact": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_query": bq, "E": E, "V": V, "X": X,
                     "kernel_ms_avg": float(kernel_ms.mean())},
        "distance_evals_per_s": E * nq / avg_kernel_s,
        "cpu_baseline": cpu,
    }
    lib.kdt_free(tree)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
