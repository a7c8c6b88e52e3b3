"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code
only): oracle/sanitize_main.cpp builds and queries trees over uniform,
duplicate-heavy and lattice inputs (widths 1-33, leafsizes 1-32, n from 1 to
300, every metric, eps, max_distance, k beyond n, one and three threads) and
walks the error paths; any memory or undefined-behaviour error aborts the run.
The reference runs its tests under valgrind memcheck the same way
(scripts/memcheck.sh)."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


def test_oracle_is_clean_under_asan_and_ubsan():
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    build = subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], capture_output=True, text=True)
    if build.returncode != 0 and "sanitize" in build.stderr and "cannot find" in build.stderr:
        pytest.skip("compiler has no sanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(ORACLE / "_san" / "oracle_sanitized")], capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "all checks passed" in run.stdout
