"""The persistent batched kernel's work order and ticket protocol, modelled on the host.

tests/sched_sim.cpp compiles the schedule functions the kernel uses (svgf_sched.hpp) with g++
and runs the ticket protocol with threads as workgroups over random ragged batches: exactly-once
execution, dependency order (a stage waits for the previous stage of its frame), no deadlock
with one worker or more workers than tickets, and counters left at zero for the next batch.
"""
from __future__ import annotations

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
