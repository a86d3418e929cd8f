"""The persistent batched kernel's work order and ticket protocol, modelled on the host.

tests/sched_sim.cpp compiles the schedule functions the kernel uses (svgf_sched.hpp) with g++
and runs the ticket protocol with threads as workgroups over random ragged batches: exactly-once
execution, dependency order (a stage waits for the previous stage of its frame), no deadlock
with one worker or more workers than tickets, and counters left at zero for the next batch.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("sched") / "sched_sim"
    subprocess.run(["g++", "-O2", "-std=c++20", "-pthread", "-Wall", "-Wextra", "-Werror", "-o", str(exe),
                    str(HERE / "sched_sim.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed,batches,workers", [(1, 40, 12), (2, 60, 4), (3, 30, 16), (4, 80, 2)])
def test_ticket_protocol(sim, seed, batches, workers):
    r = subprocess.run([str(sim), str(seed), str(batches), str(workers)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")


def test_tiles_cover_every_pixel_once(tmp_path):
    """Every stage's tiles (plain, residue pairs, tall) hold each pixel of a frame exactly once."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path / "tile_cover"
    subprocess.run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-o", str(exe),
                    str(HERE / "tile_cover.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")


def _sanitized_run(tmp_path, src, flags, args=(), env=None):
    exe = tmp_path / (src.stem + "_san")
    subprocess.run(["g++", "-O1", "-g", "-std=c++20", "-pthread", "-Wall", "-Wextra", "-Werror", *flags, "-o", str(exe),
                    str(src)], check=True)
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300,
                          env={**os.environ, **(env or {})})


def test_ticket_protocol_under_thread_sanitizer(tmp_path):
    """The ticket protocol's host model with ThreadSanitizer: its workers share the ticket, done and
    exit counters only through atomics (a plain access racing with another worker's is reported
    and fails the run)."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    r = _sanitized_run(tmp_path, HERE / "sched_sim.cpp", ["-fsanitize=thread"], (5, 12, 8),
                       {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ") and "ThreadSanitizer" not in r.stderr


def test_tiles_cover_under_address_and_ub_sanitizers(tmp_path):
    """The tile decode and counts (svgf_sched.hpp) with AddressSanitizer and UndefinedBehaviorSanitizer
    (shifts, overflow, out-of-bounds), every size and stage of the cover test."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    r = _sanitized_run(tmp_path, HERE / "tile_cover.cpp",
                       ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
