"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle's C sources (oracle/*.c) are built again with -fsanitize=address,undefined (any report
fatal) into a temporary directory, and the oracle's own test files run in a child Python process
that loads that build (SVGF_ORACLE_LIB) with the ASan runtime preloaded. An out-of-bounds access,
a use of freed memory or undefined behaviour (shift or signed overflow, misaligned access) in
the checker fails the run. CPU only.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

import pytest

HERE = pathlib.Path(__file__).resolve().parent
ORACLE = HERE.parent / "oracle"
SOURCES = ["svgf_oracle.c", "wire_oracle.c", "composite_oracle.c", "pack_oracle.c", "profiler_oracle.c"]
TEST_FILES = ["test_oracle.py", "test_golden.py", "test_pack_oracle.py", "test_composite_oracle.py",
              "test_profiler.py"]


def test_oracle_tests_under_sanitizers(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not asan or not pathlib.Path(asan).exists():
        pytest.skip("the ASan runtime is not available")
    lib = tmp_path / "libsvgf_oracle.so"
    # the oracle Makefile's flags (one IEEE rounding per float operation), plus the sanitizers
    subprocess.run(["gcc", "-O1", "-g", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-Werror", "-ffp-contract=off",
                    "-fno-fast-math", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-shared",
                    "-o", str(lib), *[str(ORACLE / s) for s in SOURCES], "-lm", "-lpthread"], check=True)
    env = {**os.environ, "SVGF_ORACLE_LIB": str(lib), "LD_PRELOAD": asan, "ASAN_OPTIONS": "detect_leaks=0",
           "UBSAN_OPTIONS": "print_stacktrace=1"}
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        *[str(HERE / t) for t in TEST_FILES]], capture_output=True, text=True, timeout=900, env=env,
                       cwd=str(HERE.parent))
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
