"""The wide FP16 state of the CANONICAL precision form (tron-grid_amd/csrc/svgf_state.hpp):
its host build against independently computed error bounds (tests/state_codec.cpp), and the
numpy decoders of the Python binding against it on the same words (CPU only)."""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

import svgf_amd as S

HERE = pathlib.Path(__file__).resolve().parent
CLANG = pathlib.Path("/opt/rocm/llvm/bin/clang++")  # binary16 arithmetic type on the host


@pytest.fixture(scope="module")
def codec_dump(tmp_path_factory):
    if not CLANG.exists():
        pytest.skip("ROCm clang++ not available")
    d = tmp_path_factory.mktemp("codec")
    exe = d / "state_codec"
    subprocess.run([str(CLANG), "-std=c++20", "-O2", "-Wall", "-Wextra", "-Werror", "-o", str(exe),
                    str(HERE / "state_codec.cpp")], check=True)
    out = subprocess.run([str(exe), str(d / "dump.bin")], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    return np.fromfile(d / "dump.bin", np.float32).reshape(-1, 3)


def test_codec_bounds(codec_dump):
    assert len(codec_dump) == 2_000_000


def test_codec_under_address_and_ub_sanitizers(tmp_path):
    """The same codec program with AddressSanitizer and UndefinedBehaviorSanitizer (shifts of the
    exponent and residual codes, conversions, out-of-bounds): it still meets its bounds."""
    if not CLANG.exists():
        pytest.skip("ROCm clang++ not available")
    exe = tmp_path / "state_codec_san"
    subprocess.run([str(CLANG), "-std=c++20", "-O1", "-g", "-Wall", "-Wextra", "-Werror",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-o", str(exe),
                    str(HERE / "state_codec.cpp")], check=True)
    out = subprocess.run([str(exe), str(tmp_path / "dump.bin")], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr


def test_numpy_decoder_matches_codec(codec_dump):
    v, h, code = codec_dump[:, 0], codec_dump[:, 1].astype(np.float16), codec_dump[:, 2].astype(np.int32)
    assert np.array_equal(h.astype(np.float32), codec_dump[:, 1])  # every h is a binary16 value
    # a history plane whose red channel and word carry the samples (length 37, g and b codes 0)
    plane = np.zeros((len(v), 4), np.float16)
    plane[:, 0] = h
    plane[:, 3] = (37 | (code << 6)).astype(np.uint16).view(np.float16)
    rgb, n = S.decode_history(plane.reshape(1, -1, 4))
    assert np.all(n == 37)
    fin = np.isfinite(h)
    dec = rgb[0, :, 0]
    # the codec's own bound, restated: |v - decoded| <= ulp(h) / 8
    e = np.floor(np.log2(np.maximum(np.abs(h[fin].astype(np.float64)), 2.0 ** -14)))
    assert np.all(np.abs(v[fin].astype(np.float64) - dec[fin]) <= np.ldexp(1.0, (e - 13).astype(np.int64)) * (1 + 1e-12))
    # the green / blue slots decode with code 0 to the half itself
    assert np.array_equal(rgb[0, :, 1], np.zeros(len(v)))


def test_numpy_moment_decoder():
    rng = np.random.default_rng(0)
    v = (np.exp(rng.uniform(-12, 10, 100000)) * rng.choice([-1, 1], 100000)).astype(np.float32)
    hi = v.astype(np.float16)
    e = (hi.view(np.uint16).astype(np.int64) >> 10) & 31
    ulp = np.ldexp(1.0, np.maximum(e, 1) - 25)
    lo = ((v.astype(np.float64) - hi.astype(np.float64)) / ulp).astype(np.float16)
    m = np.stack([hi, hi, lo, lo], axis=-1)
    d = S.decode_moments(m[None])[0]
    assert np.all(np.abs(d[:, 0] - v) <= np.abs(v) * 2.0 ** -22 + 2.0 ** -36)
    assert np.array_equal(d[:, 0], d[:, 1])
