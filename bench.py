"""Benchmark: served vs plain GPU exact k-NN on one MI355X.

Default workload (the serving configuration): 32 resident trees (n
log-uniform in [1e5, 8e6], d alternating 16 / 64, float32 standard normal
promoted to float64), Poisson arrivals of query requests on uniform trees
with batch sizes log-uniform in [1, 65536], k = 10, 200 GB HBM admission
budget.  One step = one serving run of the 240-request stream, arriving as a
Poisson process at the server's measured capacity (requests per second of a
saturated run), through the continuous-batching loop (kdtree.serve) with the
FP32 matrix-core screens and FP64 recheck.  `value` = served queries/s over
the timed steps, with p50/p99 latency (arrival to host-visible result).

Beside it, on the same requests: the plain GPU kernel (each request alone,
queries resident in HBM, one kdt_query_device call with the FP64 traversal)
on a random sample of the stream, its queries/s and the served/plain ratio;
and a byte-for-byte comparison of the two.  `roofline` is the dominant
kernel, the 64-d matrix-core screen (k_screen<32>), timed live by HIP events
around each of its launches in the timed region, against the dense bf16
matrix peak, with 2 d FLOPs per (query, point) pair as the algorithmic work.
`cpu_baseline` times the oracle (the C++ restatement of the reference's
query, threaded like its rayon path) on a bounded sample of the same requests
on this host's cores: 1 warm-up, then the median and minimum of 3 repeats.

Nested in the same line: the precision workload (n = 2^22 8-d Gaussian
mixture, 10^6 queries, k = 32: the FP64 traversal, the FP32 screen -- the tile
walk, with FP64 recheck and visit certificate -- and the traversal's own FP32
leaf screen, with the FP64 HBM roofline and the CPU ratios) and the plain GPU
kernel workload (n = 2^20 uniform 3-d points, 2^20 queries, k = 16, with its
HBM roofline).

Other workloads (--workload): c2, c3 (those lines alone), c4 (256-tree
batching), screen (the matrix-core screen alone on one 64-d tree), copy (a
1 GiB device copy that calibrates the PMC byte counters).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "kdtree-rs_amd"))

from kdtree._core import load_library  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 matrix-core peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="serve", choices=["serve", "c2", "c3", "c4", "screen", "copy"],
                    help="serve (default: the serving configuration, with the precision and plain GPU kernel "
                         "workloads nested); c2 plain GPU kernel; c3 precision; c4 multi-tree batch; screen: the "
                         "matrix-core screen alone on one 64-d tree; copy: a 1 GiB device copy that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--requests", type=int, default=240, help="serve: requests per serving run")
    ap.add_argument("--plain-seconds", type=float, default=25.0, help="serve: time budget of the plain sample")
    ap.add_argument("--no-nested", action="store_true", help="serve: skip the nested precision / plain lines")
    return ap.parse_args()


def log(msg: str):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name (lscpu, else /proc/cpuinfo), or "unknown"."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    return len(os.sched_getaffinity(0))


def cpu_timed(fn, units: float, reps: int = 3) -> dict:
    """The reference's timing protocol (benches/grid.rs:367-388): one warm-up,
    then `reps` timed repeats; value = units / median time."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    med, mn = float(np.median(ts)), float(min(ts))
    return {"value": units / med, "best": units / mn, "median_s": med, "min_s": mn, "repeats": reps}


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


def traversal_bytes(d: int, k: int, E: float, V: float, X: float) -> float:
    """Algorithmic bytes per query over the reference traversal: the query,
    the results, E leaf rows scanned (coordinates + index), V nodes visited,
    X box tests."""
    return 8 * d + 16 * k + E * (8 * d + 4) + 24 * V + 16 * d * X


def kernel_summary(times: dict) -> dict:
    return {name: {"launches": v["launches"], "ms_total": v["ms"], "ms_avg": v["ms"] / max(1, v["launches"]),
                   "busy_ms": v["busy_ms"], "work": v["work"]}
            for name, v in sorted(times.items(), key=lambda kv: -kv[1]["ms"])}


# --------------------------------------------------------------- plain GPU kernel
def c2_line(args, steps: int, warmup: int, cpu: bool) -> dict:
    """Plain GPU kernel: n = 2^20 uniform 3-d points, 2^20 queries, k = 16.
    One step = one kdt_query_device call over the resident queries, with its
    finiteness check (the reference checks every call, src/query.rs:46)."""
    import torch

    from kdtree import forest as F
    from kdtree.workloads import C2, c2_inputs
    from oracle import OracleTree

    lib = load_library()
    n, nq, d, k, ls = C2["n"], C2["nq"], C2["d"], C2["k"], C2["leafsize"]
    data_h, q_h = c2_inputs()
    data = torch.from_numpy(data_h).cuda()
    q = torch.from_numpy(q_h).cuda()
    out_d = torch.empty((nq, k), dtype=torch.float64, device="cuda")
    out_i = torch.empty((nq, k), dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = F.build(data, ls)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3

    def query():
        r = lib.kdt_query_device(tree.handle, ctypes.c_void_p(q.data_ptr()), nq, d, k, 2.0, float("inf"), 0.0, 0,
                                 ctypes.c_void_p(out_d.data_ptr()), ctypes.c_void_p(out_i.data_ptr()), sp)
        if r:
            raise RuntimeError(lib.kdt_last_error().decode())

    for _ in range(warmup):
        query()
    torch.cuda.synchronize()
    F.timing_enable(True)
    F.timing_read()
    t0 = time.perf_counter()
    for _ in range(steps):
        query()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    F.timing_enable(False)
    times = F.timing_read()
    trav = max(((n, v) for n, v in times.items() if n != "*"), key=lambda kv: kv[1]["ms"])
    kernel_ms = trav[1]["ms"] / trav[1]["launches"]

    otree = OracleTree(data_h, ls)
    od, oi, ctr = otree.query(q_h[:10_000], k=k, counters=True, threads=host_threads())
    sample = 2000
    parity = bool(np.array_equal(out_i[:sample].cpu().numpy(), oi[:sample])
                  and out_d[:sample].cpu().numpy().tobytes() == od[:sample].tobytes())
    E, V, X = (ctr / 10_000.0).tolist()
    bq = traversal_bytes(d, k, E, V, X)
    achieved = bq * nq / (kernel_ms / 1e3) / 1e9
    line = {
        "workload": "knn-uniform-n1048576-d3-q1048576-k16-leafsize32-p2",
        "value": nq * steps / wall, "unit": "queries/s", "steps": steps, "ms_per_step": wall / steps * 1e3,
        "dtype": "f64", "build_ms": build_ms, "parity_sample_bit_exact_vs_oracle": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": trav[0],
                     "kernel_ms_avg": kernel_ms, "bytes_per_query": bq, "E": E, "V": V, "X": X,
                     "what": "algorithmic bytes of the reference traversal per query (query, results, E leaf rows, "
                             "V nodes, X box tests; oracle counters on 10^4 queries) x 2^20 queries / the traversal "
                             "kernel's HIP-event time"},
        "distance_evals_per_s": E * nq / (kernel_ms / 1e3),
        "kernel_times": kernel_summary(times),
    }
    if cpu:
        threads = host_threads()
        m = 200_000
        c = cpu_timed(lambda: otree.query(q_h[:m], k=k, threads=threads), m)
        c.update(unit="queries/s", cores=threads, kind="port",
                 sample=f"{m} of the {nq} queries, k={k}, same tree, {threads} threads")
        line["cpu_baseline"] = c
        line["vs_cpu"] = line["value"] / c["value"]
    tree.free()
    return line


# --------------------------------------------------------------- precision mode
def c3_line(args, steps: int, warmup: int, cpu: bool) -> dict:
    """Precision mode: n = 2^22, d = 8, 64-component Gaussian mixture, 10^6
    queries, k = 32; the FP64 traversal versus the FP32 screens with FP64
    recheck (the tile walk, which the FP32 screen flag takes for this batch,
    and the traversal's own leaf screen), all bit-identical."""
    import torch

    from kdtree import forest as F
    from kdtree.workloads import C3, c3_inputs
    from oracle import OracleTree

    data_h, q_h = c3_inputs()
    k, ls, d = C3["k"], C3["leafsize"], C3["d"]
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
    odt, oit = F.query_device(tree, q, k, flags=F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_NO_TILE_WALK)
    F.timing_enable(True)
    F.timing_read()
    s64 = _time_steps(lambda: F.query_device(tree, q, k, od64, oi64, flags=0), steps, warmup)
    t64 = F.timing_read()
    s32 = _time_steps(lambda: F.query_device(tree, q, k, od32, oi32, flags=F.KDT_QUERY_FP32_SCREEN), steps, warmup)
    t32 = F.timing_read()
    st = _time_steps(lambda: F.query_device(tree, q, k, odt, oit, flags=F.KDT_QUERY_FP32_SCREEN
                                            | F.KDT_QUERY_NO_TILE_WALK), steps, warmup)
    tt = F.timing_read()
    F.timing_enable(False)
    same_idx = bool(torch.equal(oi64, oi32)) and bool(torch.equal(oi64, oit))
    same_bytes = bool(torch.equal(od64, od32)) and bool(torch.equal(od64, odt))
    # Rows the tile walk answers by itself (the screen-only diagnostic leaves
    # the rest unwritten).
    osd = torch.full_like(od32, float("nan"))
    osi = torch.full_like(oi32, -7)
    F.query_device(tree, q, k, osd, osi, flags=F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_SCREEN_ONLY)
    walk_answered = float((osi[:, 0] != -7).float().mean().item())
    d64, d32 = od64.cpu().numpy(), od32.cpu().numpy()
    rel = float(np.max(np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-300)))
    otree = OracleTree(data_h, ls)
    od, oi, ctr = otree.query(q_h[:10_000], k=k, counters=True, threads=host_threads())
    E_ref = float(ctr[0]) / 10_000.0
    sample = 2000
    parity = bool(np.array_equal(oi[:sample], oi64[:sample].cpu().numpy())
                  and od[:sample].tobytes() == d64[:sample].tobytes())
    E, V, X = (ctr / 10_000.0).tolist()
    bq = traversal_bytes(d, k, E, V, X)

    def roof(times):
        name, v = max(((n, v) for n, v in times.items() if n != "*"), key=lambda kv: kv[1]["ms"])
        ms = v["ms"] / v["launches"]
        ach = bq * nq / (ms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "kernel": name, "kernel_ms_avg": ms}

    line = {
        "workload": "knn-gmm64-n4194304-d8-q1000000-k32-leafsize32-p2",
        "fp64": {"value": nq / s64, "ms_per_step": s64 * 1e3, "roofline": roof(t64)},
        "fp32_screen": {"value": nq / s32, "ms_per_step": s32 * 1e3, "path": "tile walk: one tree walk per 64 "
                        "leaf-ordered queries, FP32 screen with a proven margin, FP64 recheck + visit certificate",
                        "vs_fp64": s64 / s32, "answered_by_walk": walk_answered,
                        "reference_distance_evals_per_s": E_ref * nq / s32, "kernels": kernel_summary(t32)},
        "fp32_screen_traversal": {"value": nq / st, "ms_per_step": st * 1e3, "roofline": roof(tt)},
        "unit": "queries/s", "steps": steps, "bytes_per_query": bq, "E": E, "V": V, "X": X,
        "distance_evals_per_s_fp64": E * nq / s64,
        "distance_flops_per_s_fp64": 3 * d * E * nq / s64,
        "indices_identical_fp32_vs_fp64": same_idx, "distances_identical_fp32_vs_fp64": same_bytes,
        "max_rel_distance_error_fp32_vs_fp64": rel,
        "parity_sample_bit_exact_vs_oracle": parity, "build_ms": build_ms,
    }
    if cpu:
        threads = host_threads()
        m = 40_000
        c = cpu_timed(lambda: otree.query(q_h[:m], k=k, threads=threads), m)
        c.update(unit="queries/s", cores=threads, kind="port", sample=f"{m} of the {nq} queries, {threads} threads")
        line["cpu_baseline"] = c
        line["vs_cpu_fp64"] = line["fp64"]["value"] / c["value"]
        line["vs_cpu_fp32_screen"] = line["fp32_screen"]["value"] / c["value"]
    tree.free()
    return line


# --------------------------------------------------------------- serving
def c5_points(t: int, n: int, d: int):
    import torch

    from kdtree.workloads import C5

    g = torch.Generator(device="cuda")
    g.manual_seed(C5["seed"] * 1000 + t)
    return torch.randn((n, d), generator=g, device="cuda", dtype=torch.float32).double()


def run_serve(args):
    import torch

    from kdtree import KDTree  # noqa: F401  (loads the library with torch's runtime first)
    from kdtree import forest as F
    from kdtree.serve import Request, Server, latency_stats, poisson_requests
    from kdtree.workloads import C5, c5_queries, c5_specs
    from oracle import OracleTree

    specs = c5_specs()
    dims = [d for _, d in specs]
    k = C5["k"]
    log("building 32 trees")
    trees = []
    t0 = time.perf_counter()
    for t, (n, d) in enumerate(specs):
        x = c5_points(t, n, d)
        trees.append(F.build(x, 32))
        del x
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0

    # d = 16: matrix-core screen or FP32-screened traversal, whichever is
    # faster on a sample (d = 64 always takes the matrix-core screen).
    t16 = sorted([t for t in range(len(specs)) if dims[t] == 16], key=lambda t: specs[t][0])
    t16 = t16[len(t16) // 2]
    qs = torch.from_numpy(c5_queries(10**6, 4096, 16)).cuda()
    tune = {}
    for name, fl in (("traversal_fp32_screen", F.KDT_QUERY_FP32_SCREEN),
                     ("matrix_screen", F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_MATRIX_SCREEN)):
        F.query_device(trees[t16], qs, k, flags=fl | F.KDT_QUERY_ASSUME_FINITE)
        tune[name] = _time_steps(lambda: F.query_device(trees[t16], qs, k, flags=fl | F.KDT_QUERY_ASSUME_FINITE),
                                 3, 1) * 1e3
    serve_flags = F.KDT_QUERY_FP32_SCREEN | (
        F.KDT_QUERY_MATRIX_SCREEN if tune["matrix_screen"] < tune["traversal_fp32_screen"] else 0)

    base = poisson_requests(args.requests, rate=1.0, tree_dims=dims, k=k, max_batch=C5["max_batch"],
                            seed=C5["seed"], make_queries=c5_queries)
    total_q = sum(r.queries.shape[0] for r in base)

    def fresh(rate, subset=None):
        src = base if subset is None else subset
        return [Request(r.rid, r.arrival / rate, r.tree, r.queries, r.k) for r in src]

    def serve(reqs, stats=True):
        server = Server(trees, n_streams=4, hbm_budget_bytes=C5["hbm_budget_bytes"], flags=serve_flags)
        server.run(reqs, time_limit=600)
        st = latency_stats(reqs) if stats else {}
        st.update(launches=server.launches, max_inflight=server.max_inflight)
        return st

    # Capacity: the whole stream waiting at once (also the warm-up of every
    # kernel and of the caching allocator).
    log("capacity run")
    sat = serve(fresh(1e12))
    cap_rps = sat["requests"] / sat["span_s"]
    for _ in range(args.warmup):
        serve(fresh(cap_rps), stats=False)
    torch.cuda.synchronize()

    # Timed steps: the stream at Poisson arrivals, offered at the capacity.
    log(f"timed steps at {cap_rps:.1f} requests/s")
    F.timing_enable(True)
    F.timing_read()
    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reqs = fresh(cap_rps)
        serve(reqs, stats=False)
        runs.append(reqs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    F.timing_enable(False)
    times = F.timing_read()
    lat = np.concatenate([[r.done - r.arrival for r in rs] for rs in runs]) * 1e3
    served_qps = total_q * args.steps / wall
    scr = times.get("k_screen<32>", {"launches": 0, "ms": 0.0, "work": 0.0, "busy_ms": -1.0})
    achieved = scr["work"] / (scr["ms"] / 1e3) / 1e12 if scr["ms"] else 0.0
    achieved_busy = scr["work"] / (scr["busy_ms"] / 1e3) / 1e12 if scr["busy_ms"] > 0 else None

    # Latency at lower loads.
    loads = {}
    for frac in (0.5, 0.8):
        loads[f"{frac:.1f}x_capacity"] = serve(fresh(cap_rps * frac))

    # The screen alone: 65,536 queries on the largest 64-d tree, one stream.
    big = max((t for t in range(len(specs)) if dims[t] == 64), key=lambda t: specs[t][0])
    qa = torch.from_numpy(c5_queries(2 * 10**6, 65536, 64)).cuda()
    fl = F.KDT_QUERY_FP32_SCREEN | F.KDT_QUERY_ASSUME_FINITE
    F.query_device(trees[big], qa, k, flags=fl)
    torch.cuda.synchronize()
    F.timing_enable(True)
    F.timing_read()
    alone_s = _time_steps(lambda: F.query_device(trees[big], qa, k, flags=fl), 3, 0)
    alone_t = F.timing_read()
    F.timing_enable(False)
    sa = alone_t.get("k_screen<32>", {"launches": 1, "ms": float("nan"), "work": 0.0})
    alone_tf = sa["work"] / (sa["ms"] / 1e3) / 1e12
    del qa

    # Plain GPU kernel on a random sample of the same requests, each cut to
    # its first PLAIN_CUT queries (a whole 65,536-query request on the largest
    # 64-d tree takes tens of seconds this way): each alone, queries already
    # in HBM, one kdt_query_device call (FP64 traversal).
    log("plain sample")
    PLAIN_CUT = 4096
    rng = np.random.default_rng(7)
    plain, plain_t, plain_q = [], 0.0, 0
    last = {r.rid: r for r in runs[-1]}
    for j in rng.permutation(len(base)):
        r0 = base[int(j)]
        r = Request(r0.rid, r0.arrival, r0.tree, np.ascontiguousarray(r0.queries[:PLAIN_CUT]), r0.k)
        qd = torch.from_numpy(r.queries).cuda()
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        pd, pi = F.query_device(trees[r.tree], qd, k, flags=0)
        torch.cuda.synchronize()
        dt = time.perf_counter() - s0
        plain.append((r, dt, pd.cpu().numpy(), pi.cpu().numpy()))
        plain_t += dt
        plain_q += r.queries.shape[0]
        if plain_t > args.plain_seconds and len(plain) >= 8:
            break
    plain_qps = plain_q / plain_t
    plain_lat = np.array([dt for _, dt, _, _ in plain]) * 1e3
    exact = all(np.array_equal(last[r.rid].out_i[:pi.shape[0]], pi)
                and last[r.rid].out_d[:pd.shape[0]].tobytes() == pd.tobytes() for r, _, pd, pi in plain)
    sample = [r for r, _, _, _ in plain]
    same = fresh(1e12, sample)
    served_sample = serve(same)
    exact_sample = all(np.array_equal(s.out_i, pi) and s.out_d.tobytes() == pd.tobytes()
                       for s, (_, _, pd, pi) in zip(same, plain))

    # Oracle: two requests of the sample on its smallest trees, 16 queries each.
    parity, checked = True, 0
    for r, _, pd, pi in sorted(plain, key=lambda x: specs[x[0].tree][0])[:2]:
        host = c5_points(r.tree, *specs[r.tree]).cpu().numpy()
        m = min(16, r.queries.shape[0])
        od, oi = OracleTree(host, 32).query(r.queries[:m], k=k, threads=host_threads())
        parity &= bool(np.array_equal(oi, pi[:m]) and od.tobytes() == pd[:m].tobytes())
        checked += 1

    cpu = None
    if not args.no_cpu:
        # The oracle on the first 12 requests of the plain sample, each cut to
        # 16 queries, over the trees they use (built beforehand in parallel,
        # not timed): about 10-30 s of CPU work.
        log("cpu baseline")
        threads = host_threads()
        cut = [(r.tree, np.ascontiguousarray(r.queries[:16])) for r in sample[:12]]
        need = sorted({t for t, _ in cut})

        def build_oracle(t):
            host = c5_points(t, *specs[t]).cpu().numpy()
            return t, OracleTree(host, 32)

        with ThreadPoolExecutor(max_workers=min(8, len(need))) as ex:
            otrees = dict(ex.map(build_oracle, need))
        units = sum(q.shape[0] for _, q in cut)
        cpu = cpu_timed(lambda: [otrees[t].query(q, k=k, threads=threads) for t, q in cut], units)
        cpu.update(unit="queries/s", cores=threads, kind="port", cpu_model=cpu_model(), host_cpus=os.cpu_count(),
                   sample=f"{len(cut)} requests of the plain sample on {len(need)} trees, at most 16 queries each "
                          f"({units} queries), k={k}, {threads} threads (oracle = C++ restatement of the reference "
                          "query, g++ -O3 -ffp-contract=off)")
        del otrees

    out = {
        "metric": "k-NN queries/s + p50/p99 latency, served vs plain GPU, exact indices",
        "value": served_qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp16 screen (matrix cores, f32 accumulate) + f64 recheck",
        "data": "synthetic (torch.randn float32 points, seed 5000+t; SplitMix64 Box-Muller queries)",
        "config": {"workload": "serve-32-trees-n1e5..8e6-d{16,64}-k10-batch1..65536-poisson-budget200GB",
                   "n_trees": len(specs), "n_points": int(sum(n for n, _ in specs)), "requests": args.requests,
                   "queries_per_run": int(total_q), "k": k, "offered_requests_per_s": cap_rps},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": "arrival to host-visible result, all requests of the timed steps"},
        "plain": {"queries_per_s": plain_qps, "requests": len(plain), "queries": int(plain_q),
                  "p50_ms": float(np.percentile(plain_lat, 50)), "p99_ms": float(np.percentile(plain_lat, 99)),
                  "what": "a random sample of the same requests, each cut to its first 4096 queries, each alone "
                          "through kdt_query_device (FP64 traversal, queries resident in HBM), one after another; "
                          "latency = its service time"},
        "served_vs_plain": served_qps / plain_qps,
        "served_same_sample": served_sample,
        "served_vs_plain_same_sample": served_sample["queries_per_s"] / plain_qps,
        "exact_served_vs_plain_bitwise": bool(exact and exact_sample),
        "parity_vs_oracle_sample": parity, "oracle_checked_requests": checked,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "k_screen<32>",
                     "launches": scr["launches"], "kernel_ms_avg": scr["ms"] / max(1, scr["launches"]),
                     "flops_per_launch_avg": scr["work"] / max(1, scr["launches"]),
                     "achieved_over_busy_time": achieved_busy,
                     "busy_ms": scr["busy_ms"],
                     "what": "2 d FLOPs per (query, point) pair of the 64-d matrix-core screen over the timed "
                             "serving steps / the summed HIP-event durations of its launches (each on its own "
                             "stream; launches on several streams overlap and share the device, so each one's "
                             "duration includes the sharing); achieved_over_busy_time divides by the union of "
                             "the launches' intervals instead"},
        "roofline_screen_alone": {"achieved": alone_tf, "frac": alone_tf / BF16_PEAK_TFLOPS,
                                  "kernel_ms_avg": sa["ms"] / max(1, sa["launches"]), "call_ms": alone_s * 1e3,
                                  "what": "k_screen<32> on the largest 64-d tree, 65,536 queries, one stream"},
        "kernel_times": kernel_summary(times),
        "device_busy_frac": (times["*"]["busy_ms"] / 1e3 / wall) if "*" in times and times["*"]["busy_ms"] > 0 else None,
        "served_saturation": sat, "served_at_load": loads,
        "d16_choice_ms_per_4096_queries": tune,
        "d16_matrix_screen": bool(serve_flags & F.KDT_QUERY_MATRIX_SCREEN),
        "build_s": build_s,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["vs_cpu"] = served_qps / cpu["value"]
    for tr in trees:
        tr.free()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_nested:
        log("nested: precision workload")
        out["nested"] = {"precision": c3_line(args, 5, 1, not args.no_cpu)}
        torch.cuda.empty_cache()
        log("nested: plain GPU kernel workload")
        out["nested"]["plain_gpu_kernel"] = c2_line(args, 10, 2, not args.no_cpu)
    print(json.dumps(out), flush=True)


# --------------------------------------------------------------- other workloads
def run_c2(args):
    line = c2_line(args, args.steps, args.warmup, not args.no_cpu)
    line.update(metric="k-NN queries/s (exact indices), plain GPU kernel", n_gpus=1, warmup=args.warmup,
                higher_is_better=True, scaling="weak", vs_baseline=None,
                data="synthetic (SplitMix64 uniform, seeds 2 / 1002)",
                config={"workload": line.pop("workload")})
    print(json.dumps(line), flush=True)


def run_c3(args):
    line = c3_line(args, args.steps, args.warmup, not args.no_cpu)
    line.update(metric="k-NN queries/s (exact indices), precision mode", config={"workload": line.pop("workload")})
    print(json.dumps(line), flush=True)


def run_c4(args):
    """256 trees (n log-uniform in [1e3,1e6], d in {2,3,8,16}, float32 uniform
    promoted to float64), per-tree query counts log-uniform in [1, 1e5], k in
    {1, 8, 32}; leafsize autotuned over {8,16,32,64}, lanes per query over the
    widths that hold k, and the matrix-core screen or the tile walk where they
    apply.  Shared
    persistent launches (items of 8192+ queries on their own) versus one launch
    sequence per tree versus the batch plan and its HIP graph."""
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
    for mode in ("per_tree", "persistent", "plan", "graph"):
        res[mode] = nq / _time_steps(lambda: batch.run(mode), args.steps, args.warmup)
        outs[mode] = [(od.clone(), oi.clone()) for _, _, _, od, oi in batch.items]
    modes_equal = all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
                      for m in ("per_tree", "plan", "graph") for a, b in zip(outs[m], outs["persistent"]))
    batch.run("persistent")
    torch.cuda.synchronize()
    parity = True
    for j, (data, queries, k, ls) in enumerate(host):
        od, oi = OracleTree(data, ls).query(queries, k=k, threads=host_threads())
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
        "graph_vs_persistent": res["graph"] / res["persistent"],
        "plan_bytes": int(F._lib().kdt_batch_plan_bytes(batch.plan())),
        "autotuned_leafsize_counts": leafs, "autotuned_lanes_per_query_counts": groups, "autotune_s": tune_s,
        "parity_first_8_trees_bit_exact": parity,
        "all_items_bitwise_equal_across_modes": bool(modes_equal),
        "matrix_screen_trees": sum(1 for tree, _, _ in items if tree.flags & F.KDT_QUERY_MATRIX_SCREEN),
        "tile_walk_trees": sum(1 for tree, _, _ in items if tree.flags & F.KDT_QUERY_TILE_WALK),
        "screen_slice_codes": {str(c): sum(1 for tree, _, _ in items if tree.flags & F.KDT_QUERY_MATRIX_SCREEN
                                           and (tree.flags & F.KDT_QUERY_SLICES_MASK) >> F.KDT_QUERY_SLICES_SHIFT == c)
                               for c in range(4)},
    }
    batch.close()
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
    matrix-core peak with 2 d FLOPs per (query, point) pair (the plain
    contraction) as the algorithmic work, over the k_screen kernel's own
    HIP-event time."""
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
    for _ in range(args.warmup):
        F.query_device(tree, q, k, od, oi, flags=flags)
    torch.cuda.synchronize()
    F.timing_enable(True)
    F.timing_read()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F.query_device(tree, q, k, od, oi, flags=flags)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    F.timing_enable(False)
    times = F.timing_read()
    sc = times["k_screen<32>"]
    achieved = sc["work"] / (sc["ms"] / 1e3) / 1e12
    m = 64
    otree = OracleTree(data.cpu().numpy(), 32)
    pd, pi = otree.query(q[:m].cpu().numpy(), k=k, threads=host_threads())
    parity = bool(np.array_equal(pi, oi[:m].cpu().numpy()) and pd.tobytes() == od[:m].cpu().numpy().tobytes())
    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        qs = q[:256].cpu().numpy()
        cpu = cpu_timed(lambda: otree.query(qs, k=k, threads=threads), qs.shape[0])
        cpu.update(unit="queries/s", cores=threads, kind="port", sample=f"256 of the {nq} queries, {threads} threads")
    tree.free()
    print(json.dumps({
        "metric": "k-NN queries/s (exact indices), matrix-core screen, 64-d",
        "value": nq * args.steps / wall, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "dtype": "fp16 screen (matrix cores, f32 accumulate) + f64 recheck",
        "data": "synthetic (torch.randn float32, seed 64)",
        "config": {"workload": "knn-normal-n1048576-d64-q65536-k10-leafsize32-p2"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "k_screen<32>",
                     "kernel_ms_avg": sc["ms"] / sc["launches"],
                     "what": "2 d FLOPs per (query, point) pair / the k_screen kernel's HIP-event time"},
        "kernel_times": kernel_summary(times),
        "parity_sample_bit_exact": parity, "cpu_baseline": cpu,
    }), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the workload is a single-GPU serving "
              "loop and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    {"serve": run_serve, "c2": run_c2, "c3": run_c3, "c4": run_c4, "screen": run_screen,
     "copy": run_copy}[args.workload](args)


if __name__ == "__main__":
    main()
