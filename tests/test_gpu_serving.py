"""GPU tests of the serving form: continuous batching over resident streams.

The key property (serving with uneven arrivals must not change results): every request's
output is bit-identical to denoising its stream alone, frame by frame, through the
single-frame form.
"""
from __future__ import annotations

import ctypes
import random
import time

import numpy as np
import pytest

from svgf_testutil import gpu_canonical_sequence, seq_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


SHAPES = [(96, 96), (128, 128), (33, 17), (256, 64), (70, 300), (1, 1)]


@pytest.mark.parametrize("storage", ["fp16", "fp32", "fp16_persistent", "fp16_kept_gbuffer"])
def test_server_results_equal_alone(S, storage):
    """fp16_persistent: each batch is one persistent launch, and with two lanes two such
    launches share the device. fp16_kept_gbuffer: every request keeps its G-buffer
    (SVGF_REQUEST_KEEP_GBUFFER; the inputs outlive the run), so each stream's next frame, in
    whichever lane and batch it lands, reads it in place."""
    st = S.STORAGE_FP32 if storage == "fp32" else S.STORAGE_FP16
    dt = torch.float32 if storage == "fp32" else torch.float16
    flags = S.BATCH_PERSISTENT if storage.endswith("persistent") else 0
    req_flags = S.REQUEST_KEEP_GBUFFER if storage.endswith("kept_gbuffer") else 0
    dev = torch.device("cuda")
    frames = 4
    inputs = [seq_inputs(w, h, frames, seed0=100 + 7 * i, stream=i, pan_px=0.8, disocclusion=0.1)
              for i, (w, h) in enumerate(SHAPES)]
    npx = sum(w * h for w, h in SHAPES)
    maxw = max(w for w, _ in SHAPES)
    maxh = max(h for _, h in SHAPES)
    outs = {}
    with S.Context(S.MODE_CANONICAL, st, max_width=maxw, max_height=maxh, max_streams=8, max_batch_frames=4,
                   max_batch_pixels=npx, batch_flags=flags) as ctx:
        # a partial batch waits up to 0.5 ms for more requests, so each burst of submissions is
        # packed together however fast the GPU finishes the previous batch
        with S.Server(ctx, lanes=2, max_batch_frames=4, max_batch_pixels=npx, max_wait_us=500) as srv:
            sids = [srv.stream_open(w, h) for w, h in SHAPES]
            order = [(i, f) for f in range(frames) for i in range(len(SHAPES))]
            rng = random.Random(5)
            # uploads first, so that submissions can come in tight bursts
            items = {}
            for i, f in order:
                d = inputs[i][f]
                w, h = SHAPES[i]
                items[(i, f)] = {"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
                                 "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                                 "motion": torch.from_numpy(d["motion"]).to(dev),
                                 "denoised": torch.empty((h, w, 4), dtype=dt, device=dev), "flags": req_flags}
                outs[(i, f)] = items[(i, f)]["denoised"]
            torch.cuda.synchronize()
            # random interleaving that keeps each stream's frame order; bursts of several
            # submissions (they queue behind the in-flight batches and get packed together)
            # separated by random pauses
            nxt = [0] * len(SHAPES)
            tag = 0
            while any(n < frames for n in nxt):
                for _ in range(rng.randint(1, 6)):
                    live = [k for k in range(len(SHAPES)) if nxt[k] < frames]
                    if not live:
                        break
                    i = rng.choice(live)
                    f = nxt[i]
                    nxt[i] += 1
                    srv.submit(items[(i, f)], tag)
                    tag += 1
                time.sleep(rng.random() * 1e-3)
            srv.drain()
            done = srv.poll(capacity=1024)
            assert len(done) == len(order)
            assert all(c.status == 0 for c in done)
            assert max(c.batch_frames for c in done) >= 2, "requests were never batched together"
            for c in done:
                assert c.done_ns >= c.launch_ns >= c.submit_ns
    with S.Context(S.MODE_CANONICAL, st, max_width=maxw, max_height=maxh) as sctx:
        for i in range(len(SHAPES)):
            alone = gpu_canonical_sequence(sctx, inputs[i], dt)
            for f in range(frames):
                assert np.array_equal(alone[f]["denoised"], outs[(i, f)].float().cpu().numpy()), (i, f)


def test_simulate_small(S):
    dev = torch.device("cuda")
    classes = [(96, 96), (128, 128), (320, 180)]
    class_inputs = []
    for c, (w, h) in enumerate(classes):
        ins = []
        for f in range(2):
            d = S.synth_frame(w, h, seed=f, stream=c, frame=f, pan_px=0.5)
            ins.append({"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                        "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                        "motion": torch.from_numpy(d["motion"]).to(dev)})
        class_inputs.append(ins)
    torch.cuda.synchronize()
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=320, max_height=180, max_streams=32,
                   max_batch_frames=16, max_batch_pixels=16 * 320 * 180) as ctx:
        with S.Server(ctx, lanes=2) as srv:
            r = srv.simulate(streams=24, classes=classes, class_inputs=class_inputs, rate_hz=800.0, duration_s=1.0,
                             churn_per_s=0.5, seed=3)
    print(r)
    assert r["submitted"] > 500 and r["completed"] == r["submitted"]
    assert r["streams_closed"] == r["streams_opened"]
    assert 0 < r["latency_p50_ms"] <= r["latency_p99_ms"] <= r["latency_max_ms"] < 1000
    assert r["mpx_per_s"] > 0


def test_simulate_admission_budget(S):
    """A budget that fits only some streams: the rest are rejected, nothing fails."""
    dev = torch.device("cuda")
    classes = [(128, 128)]
    d = S.synth_frame(128, 128, seed=0, frame=0)
    ins = [{"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
            "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev), "motion": torch.from_numpy(d["motion"]).to(dev)}]
    desc = S.Desc(S.MODE_CANONICAL, S.STORAGE_FP16, S.default_params(S.MODE_CANONICAL, S.STORAGE_FP16), 128, 128,
                  0, 0, 0, 0, 0, 0)
    sb = ctypes.c_uint64()
    S.check(S.lib().svgf_query_bytes(ctypes.byref(desc), 128, 128, None, ctypes.byref(sb)), "svgf_query_bytes")
    per_stream = sb.value  # history 2 x rgba16f + moments 2 x rgba16f (wide state) + G copy 2 x rgba32f
    assert per_stream >= 128 * 128 * (2 * 8 + 2 * 8 + 2 * 16)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=128, max_height=128, max_streams=16,
                   max_batch_frames=8, max_batch_pixels=8 * 128 * 128, hbm_budget_bytes=4 * per_stream + 4096) as ctx:
        with S.Server(ctx, lanes=1) as srv:
            r = srv.simulate(streams=8, classes=classes, class_inputs=[ins], rate_hz=400.0, duration_s=0.5,
                             churn_per_s=0.0, seed=1)
            assert ctx.budget_used() == 0
    assert r["streams_opened"] == 4
    assert r["rejected"] > 0 and r["completed"] == r["submitted"]
