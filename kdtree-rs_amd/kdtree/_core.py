"""kdtree._core -- the KDTree class over libkdtree_hip.so (gfx950).

Mirrors the reference's PyO3 class `kdtree._core.KDTree` (src/lib.rs:100-190):
same constructor and query signatures, the same dtype promotion and copying
(src/lib.rs:25-45), output shapes (src/lib.rs:86-98, :180), and ValueError /
TypeError messages (src/lib.rs:21-38, src/error.rs:18-51).  The build and every
query run on the GPU through the C ABI in include/kdtree_hip.h; there is no
host fallback: without the library or a device, construction raises.
ctypes releases the GIL for the duration of each call, as py.detach does
(src/lib.rs:117-118, :176-177).
"""

from __future__ import annotations

import ctypes
import operator
from pathlib import Path

import numpy as np

__all__ = ["KDTree", "library_path", "load_library"]

_HERE = Path(__file__).resolve().parent
_LIB_NAME = "libkdtree_hip.so"

# Codes of include/kdtree_hip.h: 1-11 are the reference's KDTreeError
# variants, all of which surface as ValueError (src/lib.rs:21-23).
_KDT_ERR_DEVICE = 100
_KDT_ERR_ARGUMENT = 101
KDT_QUERY_FP32_SCREEN = 0x1

_lib = None


def library_path() -> Path:
    return _HERE / _LIB_NAME


def load_library() -> ctypes.CDLL:
    """Load libkdtree_hip.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise ImportError(
            f"{path} is missing: build it with `make -C kdtree-rs_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback"
        )
    lib = ctypes.CDLL(str(path))
    vp, u64, dbl, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint32
    lib.kdt_build.argtypes = [vp, u64, u64, u64, ctypes.c_int, ctypes.POINTER(vp)]
    lib.kdt_build.restype = ctypes.c_int
    lib.kdt_build_device.argtypes = [vp, u64, u64, u64, vp, ctypes.POINTER(vp)]
    lib.kdt_build_device.restype = ctypes.c_int
    lib.kdt_query.argtypes = [vp, vp, u64, u64, u64, dbl, dbl, dbl, u32, vp, vp]
    lib.kdt_query.restype = ctypes.c_int
    lib.kdt_query_device.argtypes = [vp, vp, u64, u64, u64, dbl, dbl, dbl, u32, vp, vp, vp]
    lib.kdt_query_device.restype = ctypes.c_int
    lib.kdt_original_data.argtypes = [vp, vp]
    lib.kdt_original_data.restype = ctypes.c_int
    for name in ("kdt_ndim", "kdt_n_points", "kdt_leafsize", "kdt_n_nodes"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = u64
    lib.kdt_export.argtypes = [vp, vp, vp, vp, vp]
    lib.kdt_export.restype = ctypes.c_int
    lib.kdt_free.argtypes = [vp]
    lib.kdt_free.restype = None
    lib.kdt_last_error.argtypes = []
    lib.kdt_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _raise(lib, code: int):
    msg = lib.kdt_last_error().decode("utf-8", "replace")
    if code == _KDT_ERR_DEVICE:
        raise RuntimeError(msg)
    if code == _KDT_ERR_ARGUMENT:
        raise ValueError(msg)
    raise ValueError(msg)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


# --- argument extraction, in the manner of PyO3's (usize, f64, bool) --------
def _usize(name: str, v) -> int:
    try:
        i = operator.index(v)
    except TypeError:
        raise TypeError(
            f"argument '{name}': '{type(v).__name__}' object cannot be interpreted as an integer"
        ) from None
    if i < 0:
        raise OverflowError(f"argument '{name}': can't convert negative int to unsigned")
    if i >= 1 << 64:
        raise OverflowError(f"argument '{name}': int too big to convert")
    return i


def _f64(name: str, v) -> float:
    if isinstance(v, (str, bytes)) or not (hasattr(v, "__float__") or hasattr(v, "__index__")):
        raise TypeError(f"argument '{name}': must be real number, not {type(v).__name__}")
    return float(v)


def _boolean(name: str, v) -> bool:
    if isinstance(v, (bool, np.bool_)):
        return bool(v)
    raise TypeError(f"argument '{name}': '{type(v).__name__}' object cannot be converted to 'PyBool'")


# --- src/lib.rs:25-72 -------------------------------------------------------
def _as_f64_array(obj) -> np.ndarray:
    if not isinstance(obj, np.ndarray):
        raise TypeError(f"expected a numpy.ndarray, got {type(obj).__name__}")
    if obj.dtype == np.float64:
        return obj
    return obj.astype("float64")


class KDTree:
    """Immutable KD-tree for exact k-nearest-neighbour queries, on the GPU.

    KDTree(data, *, leafsize=32, parallel=False): `data` is an (n, m) array,
    promoted to float64 and copied; the tree keeps its own device copy.
    """

    __slots__ = ("_h", "_lib", "__weakref__")

    def __init__(self, data, *, leafsize: int = 32, parallel: bool = False) -> None:
        leafsize = _usize("leafsize", leafsize)
        parallel = _boolean("parallel", parallel)
        lib = load_library()
        arr = _as_f64_array(data)
        if arr.ndim != 2:
            raise ValueError("data must be a two-dimensional array")
        arr = np.ascontiguousarray(arr)
        n, d = arr.shape
        h = ctypes.c_void_p()
        rc = lib.kdt_build(_ptr(arr), n, d, leafsize, int(parallel), ctypes.byref(h))
        if rc:
            _raise(lib, rc)
        object.__setattr__(self, "_lib", lib)
        object.__setattr__(self, "_h", h)

    def __setattr__(self, name, value):
        raise AttributeError(f"'KDTree' object attribute '{name}' is read-only")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.kdt_free(h)
            h.value = None

    # --- properties, src/lib.rs:123-146 ---
    @property
    def ndim(self) -> int:
        return int(self._lib.kdt_ndim(self._h))

    @property
    def n_points(self) -> int:
        return int(self._lib.kdt_n_points(self._h))

    @property
    def leafsize(self) -> int:
        return int(self._lib.kdt_leafsize(self._h))

    @property
    def data(self) -> np.ndarray:
        out = np.empty((self.n_points, self.ndim), np.float64)
        rc = self._lib.kdt_original_data(self._h, _ptr(out))
        if rc:
            _raise(self._lib, rc)
        return out

    def __len__(self) -> int:
        return self.n_points

    def __repr__(self) -> str:
        return f"KDTree(n_points={self.n_points}, ndim={self.ndim}, leafsize={self.leafsize})"

    # --- src/lib.rs:161-182 ---
    def query(self, x, *, k: int = 1, p: float = 2.0, max_distance=None, eps: float = 0.0,
              parallel: bool = False, fp32_screen: bool = False):
        """k nearest neighbours of each row of x: (distances, indices).

        A 1-D x is one query and gives arrays of shape (k,); a 2-D x gives
        (n_queries, k).  Missing neighbours are (inf, n_points).
        fp32_screen=True screens leaf points in FP32 with a proven error margin
        and rechecks survivors in FP64: results are identical, traffic lower.
        """
        k = _usize("k", k)
        p = _f64("p", p)
        md = None if max_distance is None else _f64("max_distance", max_distance)
        eps = _f64("eps", eps)
        _boolean("parallel", parallel)
        arr = _as_f64_array(x)
        if arr.ndim == 1:
            single = True
            arr2 = arr.reshape(1, -1)
        elif arr.ndim == 2:
            single = False
            arr2 = arr
        else:
            raise ValueError("query must be one- or two-dimensional")
        nq, d = arr2.shape
        if d != self.ndim:
            raise ValueError(f"dimension mismatch: expected {self.ndim}, got {d}")
        arr2 = np.ascontiguousarray(arr2)
        out_d = np.empty((nq, k), np.float64)
        out_i = np.empty((nq, k), np.int64)
        flags = KDT_QUERY_FP32_SCREEN if fp32_screen else 0
        rc = self._lib.kdt_query(
            self._h, _ptr(arr2), nq, d, k, p, np.inf if md is None else md, eps, flags, _ptr(out_d), _ptr(out_i)
        )
        if rc:
            _raise(self._lib, rc)
        if single:
            return out_d.reshape(k), out_i.reshape(k)
        return out_d, out_i

    # --- introspection used by the parity tests (the reference's Tree fields) ---
    def _export(self):
        n, d = self.n_points, self.ndim
        nn = int(self._lib.kdt_n_nodes(self._h))
        data = np.empty((n, d), np.float64)
        idx = np.empty(n, np.uint32)
        nodes = np.empty(nn, NODE_DTYPE)
        boxes = np.empty((nn, 2, d), np.float64)
        rc = self._lib.kdt_export(self._h, _ptr(data), _ptr(idx), _ptr(nodes), _ptr(boxes))
        if rc:
            _raise(self._lib, rc)
        return data, idx, nodes, boxes

    @property
    def _handle(self) -> ctypes.c_void_p:
        return self._h


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
