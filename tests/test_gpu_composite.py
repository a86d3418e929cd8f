"""GPU tests of the consumer composite (svgf_composite_run) through the C ABI.

Outputs are float sums of the inputs with one rounding per operation, so the kernel must equal
the oracle (oracle/composite_oracle.c) bit for bit: all-FP32 and all-FP16 planes, per-plane
mixed formats, pitched planes, many ragged frames in one call, full 1080p / 4K frames, and the
chain the reference runs (src/mesh.slang:1007): frame t composites the denoised output of
frame t - 1.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import svgf_testutil as U

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PLANES = ("history", "emissive", "fresnel", "reflection", "reflection_colour", "reflection_base")


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _dev(a, half):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t.half() if half else t


def _items(S, scenes, half_planes, out_half):
    items = []
    for d in scenes:
        h, w = d["history"].shape[:2]
        it = {"prev_view_projection": d["m"]}
        for k in PLANES:
            it[k] = _dev(d[k], k in half_planes and k != "reflection")
        it["out"] = torch.empty((h, w, 4), dtype=torch.float16 if out_half else torch.float32, device="cuda")
        items.append(it)
    return items


def _oracle(O, d, out_half):
    return O.composite(d["m"], *[d[k] for k in PLANES], store=O.STORE_F16 if out_half else O.STORE_F32)


def _same(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


SIZES = [(61, 37), (160, 90), (1, 1), (8, 8), (33, 1), (1, 45), (1024, 3), (256, 256), (320, 180)]


@pytest.mark.parametrize("form,sizes", [("fp32", SIZES), ("fp16", SIZES), ("mixed", SIZES),
                                        ("fp16_table", SIZES * 2), ("fp32_one", SIZES[1:2])])
def test_composite_matches_oracle_bit_for_bit(S, oracle_lib, form, sizes):
    """Up to 12 frames per call travel in the kernel arguments, more in a device table: both."""
    half = {"fp32": (), "fp16": PLANES, "mixed": ("history", "fresnel", "reflection_base"), "fp16_table": PLANES,
            "fp32_one": ()}[form]
    out_half = form.startswith("fp16")
    scenes = [U.composite_scene(w, h, 100 + k, half=bool(half), coherent=k % 2 == 1) for k, (w, h) in enumerate(sizes)]
    items = _items(S, scenes, half, out_half)
    print(f"composite {form}: {len(items)} frames", flush=True)
    with S.Composite(max_frames=len(items)) as c:
        c.run(items)
        torch.cuda.synchronize()
    for (w, h), d, it in zip(sizes, scenes, items):
        want = _oracle(oracle_lib, d, out_half)
        got = it["out"].float().cpu().numpy()
        assert _same(got, want), (form, w, h, int((got != want).any(-1).sum()))


def test_composite_pitched_planes_and_no_overrun(S, oracle_lib):
    """Planes with row padding: the kernel honours each pitch and writes nothing into the output's
    padding or past its last row."""
    O = oracle_lib
    w, h = 77, 29
    d = U.composite_scene(w, h, 5)
    frame = S.CompositeFrame()
    frame.width, frame.height = w, h
    frame.prev_view_projection[:] = [float(v) for v in d["m"]]
    keep = []
    for k, pad in zip(PLANES + ("out",), (3, 1, 5, 2, 7, 4, 6)):
        half = k in ("history", "emissive") and k != "reflection"
        dt = torch.float16 if half else torch.float32
        buf = torch.full((h + 1, w + pad, 4), -3.0, dtype=dt, device="cuda")
        if k != "out":
            buf[:h, :w] = torch.from_numpy(d[k]).cuda().to(dt)
        keep.append(buf)
        esz = buf.element_size() * 4
        setattr(frame, k, S.Plane(buf.data_ptr(), (w + pad) * esz, S.FMT_RGBA16F if half else S.FMT_RGBA32F, 0))
    arr = (S.CompositeFrame * 1)(frame)
    dd = dict(d)
    dd["history"] = d["history"].astype(np.float16).astype(np.float32)
    dd["emissive"] = d["emissive"].astype(np.float16).astype(np.float32)
    with S.Composite(max_frames=1) as c:
        c.run(None, frames=arr)
        torch.cuda.synchronize()
    out = keep[-1].cpu().numpy()
    want = O.composite(dd["m"], *[dd[k] for k in PLANES])
    assert _same(out[:h, :w], want)
    assert np.all(out[:h, w:] == -3.0) and np.all(out[h:] == -3.0)


def test_composite_rejects_bad_planes(S):
    d = U.composite_scene(16, 8, 6)
    items = _items(S, [d], (), False)
    with S.Composite(max_frames=2) as c:
        bad = dict(items[0])
        bad["reflection"] = bad["reflection"].half()  # positions must be FP32
        with pytest.raises(S.SvgfError) as e:
            c.run([bad])
        assert e.value.status == 1
        bad = dict(items[0])
        bad["out"] = bad["history"]  # output over the history it gathers from
        with pytest.raises(S.SvgfError) as e:
            c.run([bad])
        assert e.value.status == 1
        with pytest.raises(S.SvgfError) as e:
            c.run(items * 3)  # over max_frames
        assert e.value.status == 6
        arr = S.Composite.frames(items)
        arr[0].prev_view_projection[5] = float("nan")
        with pytest.raises(S.SvgfError) as e:
            c.run(None, frames=arr)
        assert e.value.status == 1


@pytest.mark.parametrize("w,h,half", [(1920, 1080, True), (3840, 2160, False)])
def test_composite_full_size(S, oracle_lib, w, h, half):
    d = U.composite_scene(w, h, 21, half=half)
    items = _items(S, [d], PLANES if half else (), half)
    print(f"composite {w}x{h} half={half}", flush=True)
    with S.Composite(max_frames=1) as c:
        c.run(items)
        torch.cuda.synchronize()
    want = _oracle(oracle_lib, d, half)
    got = items[0]["out"].float().cpu().numpy()
    assert _same(got, want), int((got != want).any(-1).sum())


def test_composite_reads_last_frames_denoised_output(S, oracle_lib):
    """The reference's one-frame-late chain: resident streams denoised in a batch (FP16), then each
    stream's next-frame composite gathers from the denoised output just produced; equals the
    oracle's composite of those same denoised values."""
    O = oracle_lib
    sizes = [(96, 96), (128, 128), (320, 180), (256, 256)]
    npx = sum(w * h for w, h in sizes)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=320, max_height=256, max_streams=len(sizes),
                   max_batch_frames=len(sizes), max_batch_pixels=npx) as ctx, S.Composite(max_frames=len(sizes)) as c:
        sids = [ctx.stream_open(w, h) for w, h in sizes]
        outs = [torch.empty((h, w, 4), dtype=torch.float16, device="cuda") for w, h in sizes]
        for f in range(3):
            items = []
            for i, (w, h) in enumerate(sizes):
                dd = S.synth_frame(w, h, seed=30 * i + f, stream=i, frame=f, pan_px=0.5)
                items.append({"stream_id": sids[i], "radiance": torch.from_numpy(dd["radiance"]).cuda().half(),
                              "gbuffer": torch.from_numpy(dd["gbuffer"]).cuda(),
                              "motion": torch.from_numpy(dd["motion"]).cuda(), "denoised": outs[i]})
            ctx.denoise_batch(items)
            scenes = [U.composite_scene(w, h, 500 + 10 * f + i, half=True) for i, (w, h) in enumerate(sizes)]
            citems = _items(S, scenes, PLANES, True)
            for i in range(len(sizes)):
                citems[i]["history"] = outs[i]
            c.run(citems)
            torch.cuda.synchronize()
            for i, d in enumerate(scenes):
                d = dict(d)
                d["history"] = outs[i].float().cpu().numpy()
                want = _oracle(O, d, True)
                assert _same(citems[i]["out"].float().cpu().numpy(), want), (f, i)
