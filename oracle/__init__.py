"""ctypes loader for the oracle library (oracle/libkdtree_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

The library is a CPU restatement of the reference's build and query
(src/build.rs, src/query.rs, src/kernel.rs, src/metric.rs, src/simd.rs,
src/layout.rs); see kdtree_oracle.cpp for the per-function citations.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libkdtree_oracle.so"

NODE_DTYPE = np.dtype(
    [
        ("split_value", "<f8"),
        ("a", "<u4"),
        ("b", "<u4"),
        ("split_dim", "<u4"),
        ("is_leaf", "u1"),
        ("order_by_box", "u1"),
        ("pad", "<u2"),
    ]
)
assert NODE_DTYPE.itemsize == 24

ERRORS = {
    1: "EmptyData",
    2: "InvalidLeafsize",
    3: "InvalidK",
    4: "InvalidMetric",
    5: "InvalidEps",
    6: "InvalidMaxDistance",
    7: "NonFiniteData",
    8: "TooManyPoints",
    9: "TooManyNodes",
    10: "InvalidShape",
    11: "DimensionMismatch",
}


class OracleError(ValueError):
    def __init__(self, code: int):
        super().__init__(ERRORS.get(code, f"error {code}"))
        self.code = code


def build_library() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build_library()
        L = ctypes.CDLL(str(_LIB_PATH))
        vp, u64, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double
        L.kdo_build.argtypes = [vp, u64, u64, u64, ctypes.POINTER(vp)]
        L.kdo_build.restype = ctypes.c_int
        L.kdo_free.argtypes = [vp]
        L.kdo_free.restype = None
        L.kdo_n_nodes.argtypes = [vp]
        L.kdo_n_nodes.restype = u64
        L.kdo_count_nodes.argtypes = [u64, u64]
        L.kdo_count_nodes.restype = u64
        L.kdo_export.argtypes = [vp, vp, vp, vp, vp]
        L.kdo_export.restype = None
        L.kdo_query.argtypes = [vp, vp, u64, u64, u64, dbl, dbl, dbl, ctypes.c_int, vp, vp, vp]
        L.kdo_query.restype = ctypes.c_int
        L.kdo_point_dist.argtypes = [vp, vp, u64, dbl, u64, u64]
        L.kdo_point_dist.restype = dbl
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def count_nodes(n: int, leafsize: int) -> int:
    return int(lib().kdo_count_nodes(n, leafsize))


def point_dist(q, pt, p: float = 2.0, offset: int = 0, length: int = 1) -> float:
    q = np.ascontiguousarray(q, dtype=np.float64)
    pt = np.ascontiguousarray(pt, dtype=np.float64)
    return float(lib().kdo_point_dist(_ptr(q), _ptr(pt), q.size, p, offset, length))


class OracleTree:
    """The reference's Tree (src/tree.rs:29-36), built by the restatement."""

    def __init__(self, data: np.ndarray, leafsize: int = 32):
        data = np.ascontiguousarray(data, dtype=np.float64)
        assert data.ndim == 2
        self.n, self.d = data.shape
        self.leafsize = leafsize
        h = ctypes.c_void_p()
        rc = lib().kdo_build(_ptr(data), self.n, self.d, leafsize, ctypes.byref(h))
        if rc:
            raise OracleError(rc)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().kdo_free(h)
            self._h = None

    @property
    def n_nodes(self) -> int:
        return int(lib().kdo_n_nodes(self._h))

    def export(self):
        """(data in tree order, indices, nodes, boxes) -- the Tree fields."""
        nn = self.n_nodes
        data = np.empty((self.n, self.d), np.float64)
        idx = np.empty(self.n, np.uint32)
        nodes = np.empty(nn, NODE_DTYPE)
        boxes = np.empty((nn, 2, self.d), np.float64)
        lib().kdo_export(self._h, _ptr(data), _ptr(idx), _ptr(nodes), _ptr(boxes))
        return data, idx, nodes, boxes

    def query(self, x, k=1, p=2.0, max_distance=None, eps=0.0, threads=1, counters=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        single = x.ndim == 1
        x2 = x.reshape(1, -1) if single else x
        nq, d = x2.shape
        od = np.empty((nq, k), np.float64)
        oi = np.empty((nq, k), np.int64)
        ctr = np.zeros(3, np.uint64)
        md = np.inf if max_distance is None else float(max_distance)
        rc = lib().kdo_query(
            self._h, _ptr(x2), nq, d, k, float(p), md, float(eps), int(threads), _ptr(od), _ptr(oi), _ptr(ctr)
        )
        if rc:
            raise OracleError(rc)
        if single:
            od, oi = od[0], oi[0]
        if counters:
            return od, oi, ctr
        return od, oi


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
