"""The wire encoder's code estimate against the oracle's thresholds (CPU only).

tests/wire_code_est.c restates the estimate the encoder kernel corrects with one comparison on
each side (tron-grid_amd/csrc/wire.hip, code_of) and checks that it lies within one code of the
threshold count, the condition under which the corrected code is exact. The GPU tests then
compare the encoder with the oracle bit for bit (tests/test_gpu_wire.py).
"""
from __future__ import annotations

import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_BUILD = HERE.parent / "oracle" / "_build"


def test_code_estimate_within_one_code(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    lib = ORACLE_BUILD / "libsvgf_oracle.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(HERE.parent / "oracle")], check=True)
    exe = tmp_path / "wire_code_est"
    subprocess.run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", "-o", str(exe), str(HERE / "wire_code_est.c"),
                    f"-L{ORACLE_BUILD}", "-lsvgf_oracle", "-lm", f"-Wl,-rpath,{ORACLE_BUILD}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
