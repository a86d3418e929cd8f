"""svgf_denoise_static: the single-frame pass with no context and no allocation (its scratch is
a fixed device array of the library). Called through the C ABI, it gives svgf_denoise's results
bit for bit on two calls in a row, the second on a smaller frame, in both modes and storage
forms; frames above SVGF_STATIC_MAX_PIXELS are rejected."""
from __future__ import annotations

import ctypes

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _planes(S, mode, storage, moment_channels, d, prev=None):
    dev = torch.device("cuda")
    h, w = d["radiance"].shape[:2]
    dt = torch.float32 if storage == S.STORAGE_FP32 else torch.float16
    canon = mode == S.MODE_CANONICAL
    p = {"radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
         "denoised": torch.full((h, w, 4), 7.0, dtype=dt, device=dev),
         "moments_curr": torch.full((h, w, moment_channels), 7.0, dtype=dt, device=dev),
         "history_valid": prev is not None}
    if canon:
        p["gbuffer"] = torch.from_numpy(d["gbuffer"]).to(dev)
        p["history_curr"] = torch.full((h, w, 4), 7.0, dtype=dt, device=dev)
    if prev is not None:
        p["moments_prev"] = prev["moments_curr"]
        if canon:
            p["gbuffer_prev"] = prev["gbuffer"]
            p["motion"] = torch.from_numpy(d["motion"]).to(dev)
            p["history_prev"] = prev["history_curr"]
    return p


def _same(a, b):
    bits = torch.int16 if a.dtype == torch.float16 else torch.int32
    return torch.equal(a.view(bits), b.view(bits))


@pytest.mark.parametrize("mode,storage", [(0, 0), (0, 1), (1, 0), (1, 1)],
                         ids=["reference_fp32", "reference_fp16", "canonical_fp32", "canonical_fp16_wide"])
def test_static_matches_context(S, mode, storage):
    sizes = [(1280, 720), (320, 180)]  # the second call on a smaller frame
    params = S.default_params(mode, storage)
    for w, h in sizes:
        seq = [S.synth_frame(w, h, seed=50 + f, frame=f, pan_px=0.7, disocclusion=0.1) for f in range(2)]
        with S.Context(mode, storage, params=params, max_width=w, max_height=h) as ctx:
            ch = ctx.moment_channels
            ref_prev = got_prev = None
            for d in seq:
                a = _planes(S, mode, storage, ch, d, ref_prev)
                b = _planes(S, mode, storage, ch, d, got_prev)
                ctx.denoise(**a)
                assert S.denoise_static(mode, storage, params, **b) == 0
                torch.cuda.synchronize()
                for k in ("denoised", "moments_curr", "history_curr"):
                    if k in a:
                        assert _same(a[k], b[k]), (w, h, k)
                ref_prev, got_prev = a, b


def test_static_rejects_oversized_and_bad_arguments(S):
    w, h = 1280, 721  # one row above the maximum
    assert w * h > S.STATIC_MAX_PIXELS
    d = S.synth_frame(w, h, seed=1, frame=0)
    p = _planes(S, S.MODE_REFERENCE, S.STORAGE_FP32, 2, d)
    assert S.denoise_static(S.MODE_REFERENCE, S.STORAGE_FP32, **p) == 6  # SVGF_ERR_CAPACITY
    L = S.lib()
    params = S.default_params(S.MODE_REFERENCE, S.STORAGE_FP32)
    assert L.svgf_denoise_static(0, 0, ctypes.byref(params), None, None) == 1
    assert L.svgf_denoise_static(0, 0, None, None, None) == 1
    q = _planes(S, S.MODE_REFERENCE, S.STORAGE_FP32, 2, S.synth_frame(64, 32, seed=1, frame=0))
    f32 = S.make_frame(S.STORAGE_FP32, S.FMT_RG32F, **q)  # RGBA32F / RG32F planes for FP16 storage
    p16 = S.default_params(S.MODE_REFERENCE, S.STORAGE_FP16)
    assert L.svgf_denoise_static(0, 1, ctypes.byref(p16), ctypes.byref(f32), None) == 1
    bad = S.default_params(S.MODE_REFERENCE, S.STORAGE_FP32)
    bad.iterations = 0
    assert S.denoise_static(S.MODE_REFERENCE, S.STORAGE_FP32, bad, **q) == 1
