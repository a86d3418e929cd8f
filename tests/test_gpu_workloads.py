"""GPU tests at the sizes of the batched and serving workloads.

Batched: the 64-frame mix (resolutions drawn uniformly with seed 0 from {96x96, 128x128,
256x256, 320x180, 640x480, 1280x720}, each stream with its own seeded generator and resident
history), 3 batches, in the per-stage graph form and the persistent form: every stream's
output is bit-identical to the single-frame path, and two streams of every class match the
FP64 oracle within 1e-3 max relative error.

Serving: 256 concurrent streams, Poisson arrivals at 4000 requests/s,
5 %/s stream churn (cold history), a 64 GB HBM budget, the same resolution mix: every request
completes, and the requests of sampled streams give the same bits as their stream denoised
alone through the single-frame path.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

from svgf_testutil import moment_errors, rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
RESOLUTIONS = [(96, 96), (128, 128), (256, 256), (320, 180), (640, 480), (1280, 720)]
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _bits(t):
    return t.view(torch.int16 if t.dtype == torch.float16 else torch.int32).cpu()


def test_batched_mix64(S, oracle_lib):
    O = oracle_lib
    dev = torch.device("cuda")
    rng = np.random.default_rng(0)
    shapes = [RESOLUTIONS[i] for i in rng.integers(0, len(RESOLUTIONS), 64)]
    npx = sum(w * h for w, h in shapes)
    frames = 3
    host = [[S.synth_frame(w, h, seed=1000 * i + f, stream=i, frame=f, pan_px=0.5, disocclusion=0.05)
             for f in range(frames)] for i, (w, h) in enumerate(shapes)]
    for fr in host:
        for d in fr:
            d["radiance"] = d["radiance"].astype(np.float16).astype(np.float32)
    dev_in = [[{"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                "motion": torch.from_numpy(d["motion"]).to(dev)} for d in fr] for fr in host]
    print(f"batched mix: {len(shapes)} frames, {npx} px per batch")

    # single-frame path, stream by stream
    alone = []
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=1280, max_height=720) as ctx:
        for i, (w, h) in enumerate(shapes):
            hist = [torch.zeros((h, w, 4), dtype=torch.float16, device=dev) for _ in range(2)]
            mom = [torch.zeros((h, w, ctx.moment_channels), dtype=torch.float16, device=dev) for _ in range(2)]
            outs = []
            for f in range(frames):
                d = dev_in[i][f]
                den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
                ctx.denoise(radiance=d["radiance"], gbuffer=d["gbuffer"],
                            gbuffer_prev=dev_in[i][f - 1]["gbuffer"] if f else None, motion=d["motion"],
                            history_prev=hist[(f + 1) % 2] if f else None, history_curr=hist[f % 2],
                            moments_prev=mom[(f + 1) % 2] if f else None, moments_curr=mom[f % 2], denoised=den,
                            history_valid=f > 0)
                outs.append(den)
            torch.cuda.synchronize()
            alone.append([_bits(o) for o in outs])

    # the batched forms
    for flags in (0, S.BATCH_PERSISTENT):
        with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=1280, max_height=720, max_streams=64,
                       max_batch_frames=64, max_batch_pixels=npx, batch_flags=flags) as ctx:
            sids = [ctx.stream_open(w, h) for w, h in shapes]
            outs = [[None] * frames for _ in shapes]
            for f in range(frames):
                items = []
                for i, (w, h) in enumerate(shapes):
                    den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
                    items.append({"stream_id": sids[i], "denoised": den, **dev_in[i][f]})
                    outs[i][f] = den
                ctx.denoise_batch(items)
                assert ctx.batch_status(ctx.batch_last(), wait=True) == 0
            torch.cuda.synchronize()
            for i in range(len(shapes)):
                for f in range(frames):
                    assert torch.equal(_bits(outs[i][f]), alone[i][f]), (flags, i, f, shapes[i])
            if flags == 0:
                # two streams of every class against the FP64 oracle
                state = {}
                for i, (w, h) in enumerate(shapes):
                    st = ctx.stream_state(sids[i])
                    hist = S.device_tensor_from_plane(st.history, h, w, 4, torch.float16).cpu().numpy()
                    mom = S.device_tensor_from_plane(st.moments, h, w, 4, torch.float16).cpu().numpy()
                    state[i] = (S.decode_history(hist)[0], S.decode_moments(mom))
    op = O.canonical_params(5, compensated=True)
    worst = {}
    for cls in RESOLUTIONS:
        picks = [i for i, s in enumerate(shapes) if s == cls][:2]
        for i in picks:
            hist = mom = gp = None
            for f in range(frames):
                d = host[i][f]
                o = O.c_frame_f64(d["radiance"], d["gbuffer"], gp, d["motion"], hist, mom, op, history_valid=f > 0,
                                  threads=THREADS)
                hist, mom, gp = o["history"], o["moments"], d["gbuffer"]
                got = alone[i][f].view(torch.float16).float().numpy()
                worst["denoised"] = max(worst.get("denoised", 0.0), rel_err(got[..., :3], o["denoised"][..., :3])[0])
            h3, m2 = state[i]
            worst["history"] = max(worst.get("history", 0.0), rel_err(h3, o["history"][..., :3])[0])
            me = moment_errors(m2, o["moments"])
            worst["moments_mean"] = max(worst.get("moments_mean", 0.0), me["mean"][0])
            worst["moments_second"] = max(worst.get("moments_second", 0.0), me["second"][0])
    print("batched mix vs FP64 oracle, max rel err:", {k: f"{v:.3e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-3, (k, v)


def test_serving_256_streams(S):
    dev = torch.device("cuda")
    rate, duration, streams, churn = 4000.0, 1.5, 256, 0.05
    rng = np.random.default_rng(4)
    # two input frames per class; a stream's requests alternate them (frame 0 follows frame 1 with
    # the motion negated), so its history stays valid
    class_in = []
    for c, (w, h) in enumerate(RESOLUTIONS):
        ins = []
        for f in range(2):
            d = S.synth_frame(w, h, seed=7000 + 10 * c + f, stream=10000 + c, frame=f, pan_px=0.5, disocclusion=0.05)
            ins.append({"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                        "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                        "motion": torch.from_numpy(d["motion"] * (1.0 if f else -1.0)).to(dev)})
        class_in.append(ins)
    batch_px = 1280 * 720 * 8
    ctx = S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=1280, max_height=720, max_streams=streams + 64,
                    max_batch_frames=64, max_batch_pixels=batch_px, hbm_budget_bytes=64 << 30)
    sampled = set(range(8))  # never churned: their every request is checked
    agents = []
    submitted = completed = reopened = 0
    statuses = []
    with S.Server(ctx, lanes=2) as srv:
        def open_agent():
            c = int(rng.integers(0, len(RESOLUTIONS)))
            w, h = RESOLUTIONS[c]
            return {"id": srv.stream_open(w, h), "cls": c, "n": 0, "out": torch.empty((h, w, 4), dtype=torch.float16,
                                                                                         device=dev), "log": []}
        for _ in range(streams):
            agents.append(open_agent())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t_next = t_churn = 0.0
        t_next += rng.exponential(1.0 / rate)
        t_churn += rng.exponential(1.0 / (churn * streams))
        while True:
            now = time.perf_counter() - t0
            if now >= duration:
                break
            if t_churn <= now:
                t_churn += rng.exponential(1.0 / (churn * streams))
                k = int(rng.integers(len(sampled), streams))
                srv.stream_close(agents[k]["id"])  # waits for its queued and in-flight frames
                agents[k] = open_agent()
                reopened += 1
                continue
            if t_next > now:
                continue
            t_next += rng.exponential(1.0 / rate)
            k = int(rng.integers(0, streams))
            a = agents[k]
            w, h = RESOLUTIONS[a["cls"]]
            out = torch.empty((h, w, 4), dtype=torch.float16, device=dev) if k in sampled else a["out"]
            srv.submit({"stream_id": a["id"], "denoised": out, **class_in[a["cls"]][a["n"] % 2]}, tag=submitted)
            if k in sampled:
                a["log"].append(out)
            a["n"] += 1
            submitted += 1
            if submitted % 256 == 0:
                for c in srv.poll():
                    statuses.append(c.status)
        srv.drain()
        while True:
            got = srv.poll()
            if not got:
                break
            statuses.extend(c.status for c in got)
        completed = len(statuses)
    used = ctx.budget_used()
    ctx.close()
    print(f"serving: {submitted} requests in {duration} s ({submitted / duration:.0f}/s), {reopened} streams "
          f"reopened, {used / 2**30:.2f} GiB of stream state")
    assert submitted > 0.8 * rate * duration
    assert completed == submitted and all(s == 0 for s in statuses)
    assert reopened > 0
    # sampled streams alone, through the single-frame path
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=1280, max_height=720) as sctx:
        checked = 0
        for k in sorted(sampled):
            a = agents[k]
            w, h = RESOLUTIONS[a["cls"]]
            hist = [torch.zeros((h, w, 4), dtype=torch.float16, device=dev) for _ in range(2)]
            mom = [torch.zeros((h, w, sctx.moment_channels), dtype=torch.float16, device=dev) for _ in range(2)]
            for f, got in enumerate(a["log"]):
                d = class_in[a["cls"]][f % 2]
                den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
                sctx.denoise(radiance=d["radiance"], gbuffer=d["gbuffer"],
                             gbuffer_prev=class_in[a["cls"]][(f + 1) % 2]["gbuffer"] if f else None, motion=d["motion"],
                             history_prev=hist[(f + 1) % 2] if f else None, history_curr=hist[f % 2],
                             moments_prev=mom[(f + 1) % 2] if f else None, moments_curr=mom[f % 2], denoised=den,
                             history_valid=f > 0)
                torch.cuda.synchronize()
                assert torch.equal(_bits(den), _bits(got)), (k, f, w, h)
                checked += 1
    print(f"alone == mix on {checked} requests of {len(sampled)} streams")
    assert checked >= 8
