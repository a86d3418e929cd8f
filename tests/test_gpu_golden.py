"""The HIP path against the golden fixtures of tests/golden (FP64 expected outputs).

REFERENCE FP32 single frame within 2e-5; CANONICAL FP32 (raw moments) within 2e-4 and the
FP16 wide form within 1e-3 max relative error, in the single-frame form and in the batched form
(every fixture stream in one batch per frame); history lengths exact. Each stream carries its
own GPU state from frame to frame, as a renderer would.
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

from svgf_testutil import decode_state, gpu_canonical_sequence, moment_errors, rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


@pytest.fixture(scope="module")
def ref():
    return np.load(GOLDEN / "reference.npz")


@pytest.fixture(scope="module")
def can():
    return np.load(GOLDEN / "canonical.npz")


def _frames(can, i):
    return [{k: can[f"s{i}_{k}"][f] for k in ("radiance", "gbuffer", "motion")}
            for f in range(can[f"s{i}_radiance"].shape[0])]


def _check(got, can, i, f, form, bound, raw_moments):
    den, hist, mom = got
    assert np.isfinite(den).all(), (i, f)
    assert rel_err(den[..., :3], can[f"s{i}_{form}_denoised"][f][..., :3])[0] <= bound, (i, f, "denoised")
    want_h = can[f"s{i}_{form}_history"][f]
    assert rel_err(hist[..., :3], want_h[..., :3])[0] <= bound, (i, f, "history")
    assert np.array_equal(hist[..., 3], want_h[..., 3]), (i, f, "history length")
    want_m = can[f"s{i}_{form}_moments"][f]
    if raw_moments:
        assert rel_err(mom, want_m)[0] <= bound, (i, f, "moments")
    else:
        me = moment_errors(mom, want_m)
        assert me["mean"][0] <= bound and me["second"][0] <= bound, (i, f, "moments")


def test_reference_fp32_vs_golden(S, ref):
    dev = torch.device("cuda")
    passes = int(ref["passes"])
    p = S.default_params(S.MODE_REFERENCE, S.STORAGE_FP32)
    assert p.iterations == passes
    for i, (w, h) in enumerate(ref["shapes"].tolist()):
        mom = [torch.zeros((h, w, 2), device=dev) for _ in range(2)]
        with S.Context(S.MODE_REFERENCE, S.STORAGE_FP32, max_width=w, max_height=h) as ctx:
            for f in range(ref[f"s{i}_radiance"].shape[0]):
                rad = torch.from_numpy(ref[f"s{i}_radiance"][f]).to(dev)
                den = torch.empty((h, w, 4), device=dev)
                ctx.denoise(radiance=rad, denoised=den, moments_curr=mom[f % 2],
                            moments_prev=mom[(f + 1) % 2] if f else None, history_valid=f > 0)
                torch.cuda.synchronize()
                assert rel_err(den.cpu().numpy(), ref[f"s{i}_denoised"][f])[0] <= 2e-5, (i, f)
                assert rel_err(mom[f % 2].cpu().numpy(), ref[f"s{i}_moments"][f])[0] <= 2e-5, (i, f)


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
def test_canonical_single_frame_vs_golden(S, can, form):
    fp16 = form == "fp16_wide"
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    for i, (w, h) in enumerate(can["shapes"].tolist()):
        with S.Context(S.MODE_CANONICAL, storage, max_width=w, max_height=h) as ctx:
            assert ctx.wide == fp16
            got = gpu_canonical_sequence(ctx, _frames(can, i), dt)
        for f, g in enumerate(got):
            _check((g["denoised"], g["history"], g["moments"]), can, i, f, "comp" if fp16 else "raw",
                   1e-3 if fp16 else 2e-4, raw_moments=not fp16)


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
def test_canonical_batched_vs_golden(S, can, form):
    fp16 = form == "fp16_wide"
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    dev = torch.device("cuda")
    shapes = [tuple(s) for s in can["shapes"].tolist()]
    frames = [_frames(can, i) for i in range(len(shapes))]
    npx = sum(w * h for w, h in shapes)
    with S.Context(S.MODE_CANONICAL, storage, max_width=max(w for w, _ in shapes),
                   max_height=max(h for _, h in shapes), max_streams=len(shapes), max_batch_frames=len(shapes),
                   max_batch_pixels=npx) as ctx:
        sids = [ctx.stream_open(w, h) for w, h in shapes]
        for f in range(len(frames[0])):
            items, outs = [], []
            for i, (w, h) in enumerate(shapes):
                d = frames[i][f]
                den = torch.empty((h, w, 4), dtype=dt, device=dev)
                items.append({"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
                              "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                              "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": den})
                outs.append(den)
            ctx.denoise_batch(items)
            assert ctx.batch_status(ctx.batch_last(), wait=True) == 0
            torch.cuda.synchronize()
            for i, (w, h) in enumerate(shapes):
                st = ctx.stream_state(sids[i])
                hist, mom = decode_state(ctx, S.device_tensor_from_plane(st.history, h, w, 4, dt),
                                         S.device_tensor_from_plane(st.moments, h, w, ctx.moment_channels, dt))
                _check((outs[i].float().cpu().numpy(), hist, mom), can, i, f, "comp" if fp16 else "raw",
                       1e-3 if fp16 else 2e-4, raw_moments=not fp16)
