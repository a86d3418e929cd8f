"""The compiled hosts of include/svgf.h (tests/consumer: svgf_consumer, C++20 built by hipcc;
svgf_consumer_c, C11 built by gcc) run the pass the way TronGrid's render thread records it
(one context, one svgf_denoise per frame, moments / history / denoised ping-ponged by
current_frame, src/main.cpp:907-947, 2401-2446, 3148) on hipMallocPitch planes. Their outputs
match the oracle: REFERENCE FP32 within 1e-5 of the float restatement, REFERENCE FP16 within one
half ulp of the half-store restatement (moments bit for bit), CANONICAL FP32 within 2e-4 and
the FP16 precision form within 1e-3 of the FP64 oracle."""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from svgf_testutil import moment_errors, oracle_canonical_sequence, rel_err, seq_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
BUILD = pathlib.Path(__file__).resolve().parent / "consumer" / "_build"


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


@pytest.mark.parametrize("exe,mode,storage", [("svgf_consumer", 0, 0), ("svgf_consumer", 0, 1), ("svgf_consumer", 1, 0),
                                              ("svgf_consumer", 1, 1), ("svgf_consumer_c", 0, 0),
                                              ("svgf_consumer_c", 1, 1)])
def test_compiled_host_vs_oracle(S, oracle_lib, tmp_path, exe, mode, storage):
    O = oracle_lib
    path = BUILD / exe
    assert path.exists(), f"{path} is not built (make -C tests/consumer, run by __graft_entry__.build())"
    w, h, frames = 200, 120, 3
    fp16, canonical = storage == 1, mode == 1
    dt = np.float16 if fp16 else np.float32
    inputs = seq_inputs(w, h, frames, noise=1 if canonical and fp16 else 0, pan_px=1.37, disocclusion=0.1, half=fp16)
    for f, d in enumerate(inputs):
        d["radiance"].astype(dt).tofile(tmp_path / f"radiance_{f}.bin")
        d["gbuffer"].tofile(tmp_path / f"gbuffer_{f}.bin")
        d["motion"].tofile(tmp_path / f"motion_{f}.bin")
    out = subprocess.run([str(path), str(mode), str(storage), str(w), str(h), str(frames), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    den = [np.fromfile(tmp_path / f"denoised_{f}.bin", dt).reshape(h, w, 4).astype(np.float64) for f in range(frames)]
    if not canonical:
        om = None
        for f, d in enumerate(inputs):
            om, od, _, _ = O.r_frame_f32(d["radiance"], om, O.reference_params(3),
                                         store=O.STORE_F16 if fp16 else O.STORE_F32, history_valid=f > 0)
            mx, _ = rel_err(den[f], od)
            assert mx <= (2.0 ** -10 + 1e-6 if fp16 else 1e-5), (f, mx)
        mom = np.fromfile(tmp_path / "moments.bin", dt).reshape(h, w, 2)
        if fp16:
            assert np.array_equal(mom.astype(np.float32), om)
        else:
            assert rel_err(mom, om)[0] <= 1e-5
        return
    ref = oracle_canonical_sequence(inputs, O.canonical_params(5, compensated=fp16), threads=8)
    bound = 1e-3 if fp16 else 2e-4
    for f in range(frames):
        mx, _ = rel_err(den[f][..., :3], ref[f]["denoised"][..., :3])
        print(f"frame {f}: denoised max rel err {mx:.3e}")
        assert mx <= bound, (f, mx)
    hist = np.fromfile(tmp_path / "history.bin", dt).reshape(h, w, 4)
    if fp16:  # the wide state of the default FP16 parameters
        mom = np.fromfile(tmp_path / "moments.bin", dt).reshape(h, w, 4)
        h3, n = S.decode_history(hist)
        me = moment_errors(S.decode_moments(mom), ref[-1]["moments"])
        assert me["mean"][0] <= 1e-3 and me["second"][0] <= 1e-3, me
    else:
        h3, n = hist[..., :3], hist[..., 3]
        mom = np.fromfile(tmp_path / "moments.bin", dt).reshape(h, w, 2)
        assert rel_err(mom[..., 0], ref[-1]["moments"][..., 0])[0] <= bound
    assert rel_err(h3, ref[-1]["history"][..., :3])[0] <= bound
    assert np.array_equal(n, ref[-1]["history"][..., 3])


@pytest.mark.parametrize("storage", [0, 1], ids=["fp32", "fp16_wide"])
def test_compiled_batch_host_vs_oracle(S, oracle_lib, tmp_path, storage):
    """svgf_batch_consumer (C++20): three resident CANONICAL streams of uneven sizes, one batch per
    frame with every request keeping its G-buffer (two planes per stream, alternating), each
    batch's outcome read with svgf_batch_status, the final state read back through
    svgf_stream_state_get. Every stream against the FP64 oracle: FP32 within 2e-4, the FP16
    precision form within 1e-3; history lengths exact."""
    O = oracle_lib
    path = BUILD / "svgf_batch_consumer"
    assert path.exists(), f"{path} is not built (make -C tests/consumer, run by __graft_entry__.build())"
    fp16 = storage == 1
    dt = np.float16 if fp16 else np.float32
    shapes, frames = [(96, 64), (130, 70), (200, 120)], 4
    inputs = [seq_inputs(w, h, frames, seed0=40 + 10 * i, stream=i, noise=1 if fp16 else 0, pan_px=1.1,
                         disocclusion=0.1, half=fp16) for i, (w, h) in enumerate(shapes)]
    for i, seq in enumerate(inputs):
        for f, d in enumerate(seq):
            d["radiance"].astype(dt).tofile(tmp_path / f"s{i}_radiance_{f}.bin")
            d["gbuffer"].tofile(tmp_path / f"s{i}_gbuffer_{f}.bin")
            d["motion"].tofile(tmp_path / f"s{i}_motion_{f}.bin")
    args = [str(path), str(storage), str(frames), str(tmp_path)] + [str(v) for wh in shapes for v in wh]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    bound = 1e-3 if fp16 else 2e-4
    for i, (w, h) in enumerate(shapes):
        ref = oracle_canonical_sequence(inputs[i], O.canonical_params(5, compensated=fp16), threads=8)
        for f in range(frames):
            den = np.fromfile(tmp_path / f"s{i}_denoised_{f}.bin", dt).reshape(h, w, 4).astype(np.float64)
            assert rel_err(den[..., :3], ref[f]["denoised"][..., :3])[0] <= bound, (i, f)
        hist = np.fromfile(tmp_path / f"s{i}_history.bin", dt).reshape(h, w, 4)
        if fp16:
            h3, n = S.decode_history(hist)
            mom = S.decode_moments(np.fromfile(tmp_path / f"s{i}_moments.bin", dt).reshape(h, w, 4))
            me = moment_errors(mom, ref[-1]["moments"])
            assert me["mean"][0] <= bound and me["second"][0] <= bound, (i, me)
        else:
            h3, n = hist[..., :3], hist[..., 3]
            mom = np.fromfile(tmp_path / f"s{i}_moments.bin", dt).reshape(h, w, 2)
            assert rel_err(mom, ref[-1]["moments"])[0] <= bound, i
        assert rel_err(h3, ref[-1]["history"][..., :3])[0] <= bound, i
        assert np.array_equal(n, ref[-1]["history"][..., 3]), i


def test_compiled_agent_tick_host(S):
    """tests/consumer/svgf_agent_consumer (C++20, hipcc) runs the agent tick of the integration
    guide through the C ABI alone: per tick one batched denoise of 15 agent streams (eyes of
    8x8 .. 256x256), the composite over the denoised output, the wire encode and one push into
    the pinned ring, in sequence on one stream and pipelined (wire buffers alternating, each push
    on a copy stream once the host has seen its encode finish, the encode two ticks later
    waiting on the device for that push). Every tick's ring bytes are equal in both schedules and
    consecutive ticks differ."""
    path = BUILD / "svgf_agent_consumer"
    assert path.exists(), f"{path} is not built (make -C tests/consumer, run by __graft_entry__.build())"
    r = subprocess.run([str(path), "15", "6"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
