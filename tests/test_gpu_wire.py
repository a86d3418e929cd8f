"""GPU tests of the agent-frame wire encoder and the readback ring, through the C ABI.

The encoder's output is bytes and exactly computed floats: it must equal the oracle
(oracle/wire_oracle.c) bit for bit, on the agent sizes of docs/AI_INTERFACE.md:354-360
(8x8 .. 256x256) and on ragged sizes that end inside a 4-pixel group.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SIZES = [(8, 8), (32, 32), (64, 64), (128, 128), (256, 256), (1, 1), (3, 5), (17, 9), (255, 3), (96, 33)]


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _inputs(S, w, h, seed):
    d = S.synth_frame(w, h, seed=seed, frame=0, noise=1, pan_px=0.5)
    col = d["radiance"].copy()
    g = d["gbuffer"].copy()
    rng = np.random.default_rng(seed)
    # HDR extremes, black, and pixels without geometry
    k = max(1, w * h // 16)
    idx = rng.integers(0, w * h, k)
    col.reshape(-1, 4)[idx[: k // 2 + 1], :3] *= 400.0
    col.reshape(-1, 4)[idx[k // 2:], :3] = 0.0
    zidx = rng.integers(0, w * h, k)
    g.reshape(-1, 4)[zidx, 3] = np.float32([0.0, -2.0, np.inf, np.nan, 150.0, 0.05])[np.arange(k) % 6]
    return col, g


def _encode_all(S, sizes, dtype, params=None, seed0=0):
    dev = torch.device("cuda")
    items, host = [], []
    for k, (w, h) in enumerate(sizes):
        col, g = _inputs(S, w, h, seed0 + k)
        colour = torch.from_numpy(col).to(dev).to(dtype)
        items.append({"colour": colour, "gbuffer": torch.from_numpy(g).to(dev),
                      "rgb": torch.empty(3 * w * h + 4, dtype=torch.uint8, device=dev),
                      "depth": torch.empty(w * h + 4, dtype=torch.float32, device=dev)})
        host.append((colour.float().cpu().numpy(), g))
    with S.Wire(max_frames=len(sizes)) as wire:
        wire.encode(items, params)
        torch.cuda.synchronize()
    return items, host


@pytest.mark.parametrize("dtype", ["fp16", "fp32"])
def test_wire_matches_oracle_bit_for_bit(S, oracle_lib, dtype):
    O = oracle_lib
    dt = torch.float16 if dtype == "fp16" else torch.float32
    items, host = _encode_all(S, SIZES, dt)
    for (w, h), it, (col, g) in zip(SIZES, items, host):
        rgb, depth = O.wire_encode(col, g)
        got_rgb = it["rgb"][: 3 * w * h].cpu().numpy().reshape(h, w, 3)
        got_depth = it["depth"][: w * h].cpu().numpy().reshape(h, w)
        assert np.array_equal(got_rgb, rgb), (w, h, int((got_rgb != rgb).sum()))
        assert np.array_equal(got_depth.view(np.uint32), depth.view(np.uint32)), (w, h)


def test_wire_params_and_no_overrun(S, oracle_lib):
    """Exposure and depth planes reach the kernel; the bytes just past each frame's planes stay
    untouched (the tail groups write byte by byte)."""
    O = oracle_lib
    p = S.wire_default_params()
    p.exposure, p.depth_near, p.depth_far = 0.25, 1.0, 40.0
    dev = torch.device("cuda")
    items, host = [], []
    for k, (w, h) in enumerate(SIZES):
        col, g = _inputs(S, w, h, 100 + k)
        rgb = torch.full((3 * w * h + 8,), 0xA5, dtype=torch.uint8, device=dev)
        dep = torch.full((w * h + 8,), -7.0, dtype=torch.float32, device=dev)
        items.append({"colour": torch.from_numpy(col).to(dev), "gbuffer": torch.from_numpy(g).to(dev), "rgb": rgb,
                      "depth": dep})
        host.append((col, g))
    with S.Wire(max_frames=len(SIZES)) as wire:
        wire.encode(items, p)
        torch.cuda.synchronize()
    for (w, h), it, (col, g) in zip(SIZES, items, host):
        rgb, depth = O.wire_encode(col, g, exposure=0.25, near=1.0, far=40.0)
        r = it["rgb"].cpu().numpy()
        d = it["depth"].cpu().numpy()
        assert np.array_equal(r[: 3 * w * h].reshape(h, w, 3), rgb)
        assert np.array_equal(d[: w * h].reshape(h, w), depth)
        assert np.all(r[3 * w * h:] == 0xA5) and np.all(d[w * h:] == -7.0), (w, h)


def test_readback_ring_semantics(S):
    dev = torch.device("cuda")
    n = 4096
    srcs = [torch.full((n,), i + 1, dtype=torch.uint8, device=dev) for i in range(5)]
    torch.cuda.synchronize()
    with S.Readback(n, slots=3) as rb:
        seqs = [rb.push(s.data_ptr(), n) for s in srcs]
        assert seqs == [1, 2, 3, 4, 5]
        for q in (3, 4, 5):
            v = rb.view(q)
            assert v.size == n and np.all(v == q)
        for q in (1, 2, 9):  # overwritten, overwritten, never pushed
            with pytest.raises(S.SvgfError) as e:
                rb.view(q)
            assert e.value.status == 1
        with pytest.raises(S.SvgfError) as e:
            rb.push(srcs[0].data_ptr(), n + 1)  # larger than a slot
        assert e.value.status == 6


def test_agent_frames_end_to_end(S, oracle_lib):
    """Resident agent streams denoised in a batch (FP16), encoded to the wire in one launch and
    read back through the ring into one host buffer: each agent's bytes equal the oracle's
    encoding of the GPU's denoised output."""
    O = oracle_lib
    dev = torch.device("cuda")
    sizes = [(8, 8), (32, 32), (64, 64), (128, 128), (256, 256)] * 3
    npx = sum(w * h for w, h in sizes)
    wire_bytes = 7 * npx
    rgb_off, dep_off, off = [], [], 0
    for w, h in sizes:  # one packed wire buffer: every agent's rgb, then (16-byte aligned) depth
        rgb_off.append(off)
        off += (3 * w * h + 15) // 16 * 16
    for w, h in sizes:
        dep_off.append(off)
        off += 4 * w * h
    packed = torch.empty(off, dtype=torch.uint8, device=dev)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=256, max_height=256, max_streams=len(sizes),
                   max_batch_frames=len(sizes), max_batch_pixels=npx) as ctx, \
            S.Wire(max_frames=len(sizes)) as wire, S.Readback(off, slots=3) as rb:
        sids = [ctx.stream_open(w, h) for w, h in sizes]
        outs = [torch.empty((h, w, 4), dtype=torch.float16, device=dev) for w, h in sizes]
        gbs = []
        for f in range(2):
            items = []
            gbs = []
            for i, (w, h) in enumerate(sizes):
                d = S.synth_frame(w, h, seed=50 * i + f, stream=i, frame=f, pan_px=0.5)
                gb = torch.from_numpy(d["gbuffer"]).to(dev)
                gbs.append(gb)
                items.append({"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                              "gbuffer": gb, "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": outs[i]})
            ctx.denoise_batch(items)
        wframes = []
        for i, (w, h) in enumerate(sizes):
            wframes.append({"colour": outs[i], "gbuffer": gbs[i],
                            "rgb": packed[rgb_off[i]: rgb_off[i] + 3 * w * h],
                            "depth": packed[dep_off[i]: dep_off[i] + 4 * w * h].view(torch.float32)})
        wire.encode(wframes)
        seq = rb.push(packed.data_ptr(), off)
        host = rb.view(seq).copy()
        assert wire_bytes <= off
        for i, (w, h) in enumerate(sizes):
            rgb, depth = O.wire_encode(outs[i].float().cpu().numpy(), gbs[i].cpu().numpy())
            assert np.array_equal(host[rgb_off[i]: rgb_off[i] + 3 * w * h].reshape(h, w, 3), rgb), i
            assert np.array_equal(host[dep_off[i]: dep_off[i] + 4 * w * h].view(np.float32).reshape(h, w), depth), i


def test_pipelined_ticks_equal_sequential(S):
    """The overlapped agent tick (denoise -> composite -> encode -> readback), three schedules of
    the same 7 ticks: in sequence on one stream; pipelined (tick k encodes into wire buffer
    k % 2, the host waits for that encode once tick k + 1's denoise is enqueued and pushes the
    buffer on a copy stream, tick k + 2's encode waits on the device for that push); and
    overlapped (as pipelined, with tick k's composite and encode on a third stream beside tick
    k + 1's denoise, the denoised planes alternating by parity and tick k + 2's denoise waiting
    for tick k's composite). Every tick's ring contents are equal across the schedules, and
    consecutive ticks differ (a stale or overwritten buffer would show)."""
    from svgf_amd.scenes import composite_scene

    dev = torch.device("cuda")
    sizes = [(8, 8), (32, 32), (64, 64), (128, 128), (256, 256), (17, 9)]
    na = len(sizes)
    npx = sum(w * h for w, h in sizes)
    rgb_off, dep_off, off = [], [], 0
    for w, h in sizes:
        rgb_off.append(off)
        off += (3 * w * h + 15) // 16 * 16
    for w, h in sizes:
        dep_off.append(off)
        off += 4 * w * h
    ticks = 7
    inputs = []
    for f in range(2):
        fr = []
        for i, (w, h) in enumerate(sizes):
            d = S.synth_frame(w, h, seed=90 * i + f, stream=i, frame=f, pan_px=0.5)
            fr.append({"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                       "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                       "motion": torch.from_numpy(d["motion"] * (1.0 if f else -1.0)).to(dev)})
        inputs.append(fr)
    scenes = []
    for i, (w, h) in enumerate(sizes):
        d = composite_scene(w, h, 700 + i, half=True)
        it = {"prev_view_projection": d["m"]}
        for k in ("emissive", "fresnel", "reflection", "reflection_colour", "reflection_base"):
            t = torch.from_numpy(d[k]).to(dev)
            it[k] = t if k == "reflection" else t.half()
        scenes.append(it)

    def run(mode):
        main, side, copy = (torch.cuda.Stream(device=dev) for _ in range(3))
        if mode != "over":
            side = main
        bufs = [torch.zeros(off, dtype=torch.uint8, device=dev) for _ in range(2)]
        outs = [[torch.empty((h, w, 4), dtype=torch.float16, device=dev) for w, h in sizes] for _ in range(2)]
        comp_out = [torch.empty((h, w, 4), dtype=torch.float16, device=dev) for w, h in sizes]
        ev = {name: [torch.cuda.Event() for _ in range(2)] for name in ("den", "cmp", "enc", "rb")}
        got = {}
        with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=256, max_height=256, max_streams=na,
                       max_batch_frames=na, max_batch_pixels=npx) as ctx, \
                S.Composite(max_frames=na) as comp, S.Wire(max_frames=na) as wire, \
                S.Readback(off, slots=3) as rb:
            sids = [ctx.stream_open(w, h) for w, h in sizes]
            torch.cuda.synchronize()

            def push(k):
                b = k % 2
                ev["enc"][b].synchronize()
                got[k] = (rb.push(bufs[b].data_ptr(), off, stream=copy), None)
                ev["rb"][b].record(copy)
                if k >= 1:  # the ring holds 3 slots: read tick k - 1 before it can be overwritten
                    q, _ = got[k - 1]
                    got[k - 1] = (q, rb.view(q).copy())

            for k in range(ticks):
                b = k % 2 if mode != "seq" else 0
                ob = k % 2 if mode == "over" else 0
                if mode == "over" and k >= 2:
                    main.wait_event(ev["cmp"][ob])
                ctx.denoise_batch([{"stream_id": sids[i], "denoised": outs[ob][i], **inputs[k % 2][i]}
                                   for i in range(na)], stream=main)
                ev["den"][ob].record(main)
                side.wait_event(ev["den"][ob])
                comp.run([dict(scenes[i], history=outs[ob][i], out=comp_out[i]) for i in range(na)], stream=side)
                ev["cmp"][ob].record(side)
                wf = [{"colour": comp_out[i], "gbuffer": inputs[k % 2][i]["gbuffer"],
                       "rgb": bufs[b][rgb_off[i]: rgb_off[i] + 3 * w * h],
                       "depth": bufs[b][dep_off[i]: dep_off[i] + 4 * w * h].view(torch.float32)}
                      for i, (w, h) in enumerate(sizes)]
                if mode == "seq":
                    wire.encode(wf, stream=main)
                    q = rb.push(bufs[b].data_ptr(), off, stream=main)
                    got[k] = (q, rb.view(q).copy())
                    continue
                if k >= 1:
                    push(k - 1)
                if k >= 2:
                    side.wait_event(ev["rb"][b])
                wire.encode(wf, stream=side)
                ev["enc"][b].record(side)
            if mode != "seq":
                push(ticks - 1)
                q, _ = got[ticks - 1]
                got[ticks - 1] = (q, rb.view(q).copy())
            torch.cuda.synchronize()
        return [got[k][1] for k in range(ticks)]

    seq = run("seq")
    for mode in ("pipe", "over"):
        other = run(mode)
        for k in range(ticks):
            assert np.array_equal(seq[k], other[k]), (mode, k)
    for k in range(1, ticks):
        assert not np.array_equal(seq[k][: 7 * npx], seq[k - 1][: 7 * npx]), k


def test_agent_tick_at_256_streams(S, oracle_lib):
    """One agent tick at the bench's size: 256 resident streams at eye sizes drawn from
    docs/AI_INTERFACE.md:354-360 (8x8 .. 256x256, 4.6 Mpx), a second frame of each stream denoised
    in one batch, the composite over the denoised output in one launch, the wire encode in one
    launch, and one push into the ring. For 16 sampled agents (every size class) the composite
    equals the composite oracle on the GPU's denoised output and the ring's bytes equal the wire
    oracle's encoding of the GPU's composited output, bit for bit."""
    from svgf_amd.scenes import composite_scene

    O = oracle_lib
    dev = torch.device("cuda")
    agent_sizes = [(8, 8), (32, 32), (64, 64), (128, 128), (256, 256)]
    rng = np.random.default_rng(2)
    n = 256
    sizes = [agent_sizes[i] for i in rng.integers(0, len(agent_sizes), n)]
    npx = sum(w * h for w, h in sizes)
    rgb_off, dep_off, off = [], [], 0
    for w, h in sizes:
        rgb_off.append(off)
        off += (3 * w * h + 15) // 16 * 16
    for w, h in sizes:
        dep_off.append(off)
        off += 4 * w * h
    packed = torch.zeros(off, dtype=torch.uint8, device=dev)
    cplanes = ("emissive", "fresnel", "reflection", "reflection_colour", "reflection_base")
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=256, max_height=256, max_streams=n,
                   max_batch_frames=n, max_batch_pixels=npx) as ctx, \
            S.Composite(max_frames=n) as comp, S.Wire(max_frames=n) as wire, S.Readback(off, slots=3) as rb:
        sids = [ctx.stream_open(w, h) for w, h in sizes]
        outs = [torch.empty((h, w, 4), dtype=torch.float16, device=dev) for w, h in sizes]
        gbs = None
        for f in range(2):
            items, gbs = [], []
            for i, (w, h) in enumerate(sizes):
                d = S.synth_frame(w, h, seed=11 * i + f, stream=i, frame=f, pan_px=0.5, disocclusion=0.05)
                gb = torch.from_numpy(d["gbuffer"]).to(dev)
                gbs.append(gb)
                items.append({"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                              "gbuffer": gb, "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": outs[i]})
            ctx.denoise_batch(items)
        citems = []
        for i, (w, h) in enumerate(sizes):
            d = composite_scene(w, h, 5000 + i, half=True)
            it = {"prev_view_projection": d["m"], "history": outs[i],
                  "out": torch.empty((h, w, 4), dtype=torch.float16, device=dev)}
            for k in cplanes:
                t = torch.from_numpy(d[k]).to(dev)
                it[k] = t if k == "reflection" else t.half()
            citems.append(it)
        comp.run(citems)
        wire.encode([{"colour": citems[i]["out"], "gbuffer": gbs[i],
                      "rgb": packed[rgb_off[i]: rgb_off[i] + 3 * w * h],
                      "depth": packed[dep_off[i]: dep_off[i] + 4 * w * h].view(torch.float32)}
                     for i, (w, h) in enumerate(sizes)])
        host = rb.view(rb.push(packed.data_ptr(), off)).copy()
        torch.cuda.synchronize()
        by_class = {}
        for i, wh in enumerate(sizes):
            by_class.setdefault(wh, []).append(i)
        sample = sorted({i for idx in by_class.values() for i in idx[:3]} | {0, n - 1})[:16]
        assert {sizes[i] for i in sample} == set(agent_sizes)
        for i in sample:
            w, h = sizes[i]
            it = citems[i]
            want = O.composite(it["prev_view_projection"], *[it[k].float().cpu().numpy() for k in
                               ("history", "emissive", "fresnel", "reflection", "reflection_colour",
                                "reflection_base")], store=O.STORE_F16)
            got = it["out"].float().cpu().numpy()
            assert np.array_equal(got, want), i
            rgb, depth = O.wire_encode(got, gbs[i].cpu().numpy())
            assert np.array_equal(host[rgb_off[i]: rgb_off[i] + 3 * w * h].reshape(h, w, 3), rgb), i
            assert np.array_equal(host[dep_off[i]: dep_off[i] + 4 * w * h].view(np.float32).reshape(h, w), depth), i
