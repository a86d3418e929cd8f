"""GPU tests of the producer pack (svgf_pack_run) through the C ABI.

Every output is a fixed sequence of float operations on the inputs, so the kernel must equal
the oracle (oracle/pack_oracle.c) bit for bit: FP32 and FP16 lighting / radiance planes,
ragged frames in one call (kernel-argument and device-table descriptor paths), pitched planes
with untouched padding, full 1080p / 4K frames, and the chain into the denoiser: packing on the
GPU and then denoising gives the same result as denoising the oracle's planes.
"""
from __future__ import annotations

import numpy as np
import pytest

from svgf_amd.scenes import pack_scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LIGHT = ("indirect", "direct_diffuse", "direct_specular")


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _items(scenes, half):
    items = []
    for d in scenes:
        h, w = d["position"].shape[:2]
        it = {"view_projection": d["vp"], "prev_view_projection": d["pvp"],
              "position": torch.from_numpy(d["position"]).cuda(),
              "normal_oct": torch.from_numpy(d["normal_oct"].view(np.int32)).cuda()}
        for k in LIGHT:
            t = torch.from_numpy(d[k]).cuda()
            it[k] = t.half() if half else t
        it["radiance"] = torch.full((h, w, 4), 7.0, dtype=torch.float16 if half else torch.float32, device="cuda")
        it["gbuffer"] = torch.full((h, w, 4), 7.0, dtype=torch.float32, device="cuda")
        it["motion"] = torch.full((h, w, 2), 7.0, dtype=torch.float32, device="cuda")
        items.append(it)
    return items


def _oracle(O, d, half):
    return O.pack(d["vp"], d["pvp"], d["position"], d["normal_oct"], *[d[k] for k in LIGHT],
                  store=O.STORE_F16 if half else O.STORE_F32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


SIZES = [(61, 37), (160, 90), (1, 1), (8, 8), (33, 1), (1, 45), (1024, 3), (256, 256), (320, 180)]


@pytest.mark.parametrize("half,sizes", [(False, SIZES), (True, SIZES), (True, SIZES * 2), (False, SIZES[7:8])])
def test_pack_matches_oracle_bit_for_bit(S, oracle_lib, half, sizes):
    """Up to 10 frames per call travel in the kernel arguments, more in a device table."""
    scenes = [pack_scene(w, h, 200 + k, half=half, move=0.02 if k % 2 else 0.3) for k, (w, h) in enumerate(sizes)]
    items = _items(scenes, half)
    print(f"pack half={half}: {len(items)} frames", flush=True)
    with S.Pack(max_frames=len(items)) as p:
        p.run(items)
        torch.cuda.synchronize()
    for (w, h), d, it in zip(sizes, scenes, items):
        rad, g, mv = _oracle(oracle_lib, d, half)
        for name, want in (("radiance", rad), ("gbuffer", g), ("motion", mv)):
            got = it[name].float().cpu().numpy()
            bad = _bits(got) != _bits(want)
            assert not bad.any(), (half, w, h, name, int(bad.sum()))


def test_pack_pitched_planes_and_no_overrun(S, oracle_lib):
    w, h = 77, 29
    d = pack_scene(w, h, 9, half=True)
    f = S.PackFrame()
    f.width, f.height = w, h
    f.view_projection[:] = [float(v) for v in d["vp"]]
    f.prev_view_projection[:] = [float(v) for v in d["pvp"]]
    bufs = {}
    spec = {"position": (4, torch.float32, S.FMT_RGBA32F, 3), "normal_oct": (1, torch.int32, S.FMT_R32UI, 5),
            "indirect": (4, torch.float16, S.FMT_RGBA16F, 1), "direct_diffuse": (4, torch.float32, S.FMT_RGBA32F, 2),
            "direct_specular": (4, torch.float16, S.FMT_RGBA16F, 7), "radiance": (4, torch.float16, S.FMT_RGBA16F, 4),
            "gbuffer": (4, torch.float32, S.FMT_RGBA32F, 6), "motion": (2, torch.float32, S.FMT_RG32F, 3)}
    for name, (c, dt, fmt, pad) in spec.items():
        buf = torch.full((h + 1, w + pad, c), -3, dtype=dt, device="cuda")
        if name in ("position", "indirect", "direct_diffuse", "direct_specular"):
            buf[:h, :w] = torch.from_numpy(d[name]).cuda().to(dt)
        elif name == "normal_oct":
            buf[:h, :w, 0] = torch.from_numpy(d[name].view(np.int32)).cuda()
        bufs[name] = buf
        setattr(f, name, S.Plane(buf.data_ptr(), (w + pad) * c * buf.element_size(), fmt, 0))
    with S.Pack(max_frames=1) as p:
        p.run(None, frames=(S.PackFrame * 1)(f))
        torch.cuda.synchronize()
    rad, g, mv = _oracle(oracle_lib, d, True)
    for name, want in (("radiance", rad), ("gbuffer", g), ("motion", mv)):
        out = bufs[name].float().cpu().numpy()
        assert np.array_equal(_bits(out[:h, :w]), _bits(want)), name
        assert np.all(out[:h, w:] == -3) and np.all(out[h:] == -3), name


def test_pack_rejects_bad_planes(S):
    d = pack_scene(16, 8, 10)
    items = _items([d], False)
    with S.Pack(max_frames=2) as p:
        bad = dict(items[0])
        bad["gbuffer"] = bad["position"]  # output over an input
        with pytest.raises(S.SvgfError) as e:
            p.run([bad])
        assert e.value.status == 1
        arr = S.Pack.frames(items)
        arr[0].normal_oct.format = S.FMT_RGBA16F
        with pytest.raises(S.SvgfError) as e:
            p.run(None, frames=arr)
        assert e.value.status == 1
        arr = S.Pack.frames(items)
        arr[0].prev_view_projection[0] = float("inf")
        with pytest.raises(S.SvgfError) as e:
            p.run(None, frames=arr)
        assert e.value.status == 1
        with pytest.raises(S.SvgfError) as e:
            p.run(items * 3)
        assert e.value.status == 6


@pytest.mark.parametrize("w,h,half", [(1920, 1080, True), (3840, 2160, False)])
def test_pack_full_size(S, oracle_lib, w, h, half):
    d = pack_scene(w, h, 11, half=half)
    items = _items([d], half)
    print(f"pack {w}x{h} half={half}", flush=True)
    with S.Pack(max_frames=1) as p:
        p.run(items)
        torch.cuda.synchronize()
    rad, g, mv = _oracle(oracle_lib, d, half)
    for name, want in (("radiance", rad), ("gbuffer", g), ("motion", mv)):
        got = items[0][name].float().cpu().numpy()
        bad = _bits(got) != _bits(want)
        assert not bad.any(), (name, int(bad.sum()))


def test_pack_feeds_the_denoiser(S, oracle_lib):
    """Pack on the GPU, then denoise (CANONICAL, FP16, batched) three frames of two streams; the
    outputs equal those of the same denoiser fed the oracle's packed planes."""
    sizes = [(160, 90), (96, 64)]
    npx = sum(w * h for w, h in sizes)
    outs = {}
    for source in ("gpu", "oracle"):
        with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=160, max_height=90, max_streams=2,
                       max_batch_frames=2, max_batch_pixels=npx) as ctx, S.Pack(max_frames=2) as p:
            sids = [ctx.stream_open(w, h) for w, h in sizes]
            res = []
            for f in range(3):
                scenes = [pack_scene(w, h, 40 * i + f, half=True) for i, (w, h) in enumerate(sizes)]
                if source == "gpu":
                    items = _items(scenes, True)
                    p.run(items)
                    planes = [(it["radiance"], it["gbuffer"], it["motion"]) for it in items]
                else:
                    planes = []
                    for d in scenes:
                        rad, g, mv = _oracle(oracle_lib, d, True)
                        planes.append((torch.from_numpy(rad).cuda().half(), torch.from_numpy(g).cuda(),
                                       torch.from_numpy(mv).cuda()))
                den = [torch.empty((h, w, 4), dtype=torch.float16, device="cuda") for w, h in sizes]
                ctx.denoise_batch([{"stream_id": sids[i], "radiance": planes[i][0], "gbuffer": planes[i][1],
                                    "motion": planes[i][2], "denoised": den[i]} for i in range(len(sizes))])
                torch.cuda.synchronize()
                res.append([t.float().cpu().numpy() for t in den])
            outs[source] = res
    for f in range(3):
        for i in range(len(sizes)):
            assert np.array_equal(_bits(outs["gpu"][f][i]), _bits(outs["oracle"][f][i])), (f, i)
