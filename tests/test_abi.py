"""The C ABI loads, exports every declared symbol, and validates like the
reference before touching the device (no GPU needed)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from kdtree import KDTree
from kdtree._core import library_path, load_library

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "kdtree_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t|const char\*)\s+(kdt_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(str(library_path()))
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import kdtree._core as core

    monkeypatch.setattr(core, "_lib", None)
    monkeypatch.setattr(core, "_HERE", tmp_path)
    with pytest.raises(ImportError):
        core.load_library()


# Errors the reference raises before any work (src/lib.rs, src/build.rs:25-43,
# src/query.rs:38-57); texts from src/error.rs:18-51.
def test_type_error_for_non_arrays():
    with pytest.raises(TypeError, match="expected a numpy.ndarray, got list"):
        KDTree([[0.0, 1.0]])


def test_data_shape_error():
    with pytest.raises(ValueError, match="data must be a two-dimensional array"):
        KDTree(np.zeros(3))


def test_leafsize_zero():
    with pytest.raises(ValueError, match="^leafsize must be greater than zero$"):
        KDTree(np.zeros((3, 2)), leafsize=0)


@pytest.mark.parametrize("shape", [(0, 3), (4, 0)])
def test_empty_data(shape):
    with pytest.raises(ValueError, match="^data must contain at least one point$"):
        KDTree(np.zeros(shape))


def test_leafsize_checked_before_emptiness():
    with pytest.raises(ValueError, match="leafsize"):
        KDTree(np.zeros((0, 3)), leafsize=0)


def test_negative_leafsize_overflow():
    with pytest.raises(OverflowError):
        KDTree(np.zeros((3, 2)), leafsize=-1)


def _lib_build_rc(n, d, leafsize):
    lib = load_library()
    h = ctypes.c_void_p()
    buf = np.zeros(1)
    rc = lib.kdt_build(ctypes.c_void_p(buf.ctypes.data), n, d, leafsize, 0, ctypes.byref(h))
    return rc, lib.kdt_last_error().decode()


def test_too_many_points_message():
    rc, msg = _lib_build_rc(1 << 32, 1, 32)
    assert rc == 8
    assert msg == "n_points must fit in a 32-bit index, got 4294967296"


def test_too_many_nodes_message():
    # leafsize 1 over 2^31 + 1 points needs 2^32 + 1 nodes.
    rc, msg = _lib_build_rc((1 << 31) + 1, 1, 1)
    assert rc == 9
    assert msg == (
        "tree would need 4294967297 nodes, which exceeds the 32-bit node index; increase leafsize"
    )


def test_batch_item_layout_matches_header(tmp_path):
    # The Python mirror of kdt_batch_item (kdtree.forest.BatchItem) against the
    # C compiler's layout of the header's struct: size and every field offset.
    import shutil
    import subprocess

    from kdtree.forest import BatchItem

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f for f, _ in BatchItem._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "kdtree_hip.h"\nint main(void) {\n'
        '  printf("%zu\\n", sizeof(kdt_batch_item));\n'
        + "".join(f'  printf("%zu\\n", offsetof(kdt_batch_item, {f}));\n' for f in fields)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(BatchItem)
    assert out[1:] == [getattr(BatchItem, f).offset for f in fields]


def test_bench_refuses_more_than_one_gpu():
    # The benchmark is a single-GPU workload: --gpus above 1 exits non-zero
    # with a message, before touching any device.
    import subprocess
    import sys

    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0
    assert "one GPU only" in (r.stdout + r.stderr)
