"""The hosts of include/svgf.h compile with -Werror against the header and link against the
library: single frame in C++20 (hipcc) and C11 (gcc -std=c11 -pedantic), resident streams and
batches in C++20 (CPU only; tests/test_gpu_consumer.py runs them)."""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent


def test_consumers_compile(tmp_path):
    if not pathlib.Path("/opt/rocm/bin/hipcc").exists() or not shutil.which("gcc"):
        pytest.skip("hipcc / gcc not available")
    subprocess.run(["make", "-s", "-B", "-C", str(HERE / "consumer"), f"OUT={tmp_path}"], check=True)
    for exe in ("svgf_consumer", "svgf_consumer_c", "svgf_batch_consumer"):
        assert (tmp_path / exe).exists()
        # usage error path without a GPU: wrong argument count
        # (built outside the repository, the relative runpath does not reach the library)
        env = dict(os.environ, LD_LIBRARY_PATH=str(HERE.parent / "tron-grid_amd" / "lib"))
        out = subprocess.run([str(tmp_path / exe)], capture_output=True, text=True, env=env)
        assert out.returncode == 2 and "usage" in out.stderr
