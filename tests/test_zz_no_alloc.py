"""kdt_query_device_noalloc: the device query with its scratch in a fixed
device array, called through the C ABI, gives byte for byte what
kdt_query_device gives -- on two calls in a row, the second smaller, through
the leaf-order sort (8192 queries or more), through the matrix-core screen and
through the tile walk (or the traversal screen where its buffers do not fit)
-- rejects a query whose scratch would not fit before any launch, and checks
finiteness first."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def lib_with_noalloc():
    from kdtree import forest as F

    lib = F._lib()
    vp, u64, dbl, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint32
    lib.kdt_query_device_noalloc.argtypes = [vp, vp, u64, u64, u64, dbl, dbl, dbl, u32, vp, vp, vp]
    lib.kdt_query_device_noalloc.restype = ctypes.c_int
    return lib


def call(lib, fn, tree, q, k, flags, out_d, out_i):
    import torch

    return getattr(lib, fn)(tree.handle, ctypes.c_void_p(q.data_ptr()), q.shape[0], q.shape[1], k, 2.0,
                            float("inf"), 0.0, flags, ctypes.c_void_p(out_d.data_ptr()),
                            ctypes.c_void_p(out_i.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def compare(lib, tree, q, k, flags):
    import torch

    outs = []
    for fn in ("kdt_query_device", "kdt_query_device_noalloc"):
        od = torch.full((q.shape[0], k), -1.0, dtype=torch.float64, device="cuda")
        oi = torch.full((q.shape[0], k), -1, dtype=torch.int64, device="cuda")
        rc = call(lib, fn, tree, q, k, flags, od, oi)
        assert rc == 0, (fn, lib.kdt_last_error().decode())
        torch.cuda.synchronize()
        outs.append((od.cpu().numpy(), oi.cpu().numpy()))
    (a, b), (c, d) = outs
    assert np.array_equal(b, d) and a.tobytes() == c.tobytes()
    assert np.all(d >= 0)


@pytest.mark.parametrize("flags", [0, 0x2])  # with and without the finiteness check
def test_noalloc_matches_device_query_leaf_order(flags):
    import torch

    from kdtree import forest as F

    lib = lib_with_noalloc()
    rng = np.random.default_rng(71)
    tree = F.build(rng.uniform(size=(200_000, 3)), 32)
    try:
        for nq in (50_000, 9_000, 700):  # leaf-order sort twice, then none
            print(f"3-d, {nq} queries, flags {flags}", flush=True)
            compare(lib, tree, torch.from_numpy(rng.uniform(size=(nq, 3))).cuda(), 16, flags)
    finally:
        torch.cuda.synchronize()
        tree.free()


def test_noalloc_matches_device_query_matrix_screen():
    import torch

    from kdtree import forest as F

    lib = lib_with_noalloc()
    rng = np.random.default_rng(72)
    data = rng.standard_normal((60_000, 64)).astype(np.float32).astype(np.float64)
    tree = F.build(data, 32)
    try:
        for nq in (20_000, 3_000):
            print(f"64-d screen, {nq} queries", flush=True)
            q = torch.from_numpy(rng.standard_normal((nq, 64)).astype(np.float32).astype(np.float64)).cuda()
            compare(lib, tree, q, 10, F.KDT_QUERY_FP32_SCREEN)
    finally:
        torch.cuda.synchronize()
        tree.free()


def test_noalloc_matches_device_query_tile_walk():
    """8-d FP32 screen: 100,000 queries take the tile walk in kdt_query_device,
    whose per-wavefront buffers exceed the fixed array, so the no-allocation
    call runs the traversal's FP32 screen instead; 9,000 queries fit and run
    the walk in both.  Either way the bytes agree."""
    import torch

    from kdtree import forest as F

    lib = lib_with_noalloc()
    rng = np.random.default_rng(74)
    tree = F.build(rng.standard_normal((150_000, 8)), 32)
    try:
        for nq in (100_000, 9_000):
            print(f"8-d walk, {nq} queries", flush=True)
            compare(lib, tree, torch.from_numpy(rng.standard_normal((nq, 8))).cuda(), 16, F.KDT_QUERY_FP32_SCREEN)
    finally:
        torch.cuda.synchronize()
        tree.free()


def test_noalloc_rejects_oversized_and_non_finite_before_launch():
    import torch

    from kdtree import forest as F

    lib = lib_with_noalloc()
    rng = np.random.default_rng(73)
    tree = F.build(rng.standard_normal((5_000, 64)).astype(np.float32).astype(np.float64), 32)
    try:
        q = torch.zeros((200_000, 64), dtype=torch.float64, device="cuda")
        od = torch.full((200_000, 10), 3.0, dtype=torch.float64, device="cuda")
        oi = torch.full((200_000, 10), 3, dtype=torch.int64, device="cuda")
        rc = call(lib, "kdt_query_device_noalloc", tree, q, 10, F.KDT_QUERY_FP32_SCREEN, od, oi)
        assert rc == 101 and "bytes of scratch" in lib.kdt_last_error().decode()
        q2 = torch.zeros((4, 64), dtype=torch.float64, device="cuda")
        q2[2, 5] = float("nan")
        od2 = torch.full((4, 10), 3.0, dtype=torch.float64, device="cuda")
        oi2 = torch.full((4, 10), 3, dtype=torch.int64, device="cuda")
        rc = call(lib, "kdt_query_device_noalloc", tree, q2, 10, 0, od2, oi2)
        assert rc == 7 and lib.kdt_last_error().decode() == "all coordinates must be finite"
        torch.cuda.synchronize()
        assert bool(torch.all(od == 3.0)) and bool(torch.all(oi2 == 3)) and bool(torch.all(od2 == 3.0))
    finally:
        torch.cuda.synchronize()
        tree.free()
