"""The batched form with SVGF_REQUEST_KEEP_GBUFFER.

A request with the flag leaves its G-buffer with the caller until the stream's next frame has
completed; that frame reads it in place as its previous G-buffer and the temporal stage saves no
copy. The tests run the same multi-stream sequence with copied and with kept G-buffers (the
flag switched on and off between frames) and require the same bits in every output and in the
resident state. Each kept G-buffer is filled with NaN once the frame after it has completed, so
a frame that read an older G-buffer than its predecessor's would not match.
"""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPES = [(96, 64), (130, 70), (64, 64), (200, 120)]
FRAMES = 6
KEEP = [1, 1, 0, 1, 0, 1]  # per frame: kept or copied G-buffer


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _bits(t):
    return t.view(torch.int16 if t.dtype == torch.float16 else torch.int32).cpu()


def _inputs(S, dtype):
    dev = torch.device("cuda")
    out = []
    for i, (w, h) in enumerate(SHAPES):
        fr = []
        for f in range(FRAMES):
            d = S.synth_frame(w, h, seed=300 + 10 * i + f, stream=i, frame=f, pan_px=0.7, disocclusion=0.05)
            fr.append({"radiance": torch.from_numpy(d["radiance"]).to(dev).to(dtype),
                       "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                       "motion": torch.from_numpy(d["motion"]).to(dev)})
        out.append(fr)
    return out


def _run(S, storage, launch, dtype, keep):
    """Denoised frames (bits) and the final resident state of every stream."""
    ins = _inputs(S, dtype)
    npx = sum(w * h for w, h in SHAPES)
    dev = torch.device("cuda")
    outs = [[None] * FRAMES for _ in SHAPES]
    with S.Context(S.MODE_CANONICAL, storage, max_width=256, max_height=128, max_streams=len(SHAPES),
                   max_batch_frames=len(SHAPES), max_batch_pixels=npx, batch_flags=launch) as ctx:
        sids = [ctx.stream_open(w, h) for w, h in SHAPES]
        for f in range(FRAMES):
            flags = S.REQUEST_KEEP_GBUFFER if keep and KEEP[f] else 0
            items = []
            for i, (w, h) in enumerate(SHAPES):
                den = torch.empty((h, w, 4), dtype=dtype, device=dev)
                items.append({"stream_id": sids[i], "denoised": den, "flags": flags, **ins[i][f]})
                outs[i][f] = den
            ctx.denoise_batch(items)
            assert ctx.batch_status(ctx.batch_last(), wait=True) == 0
            torch.cuda.synchronize()
            for i, (w, h) in enumerate(SHAPES):
                st = ctx.stream_state(sids[i])
                g = ins[i][f]["gbuffer"]
                if flags:
                    # the next frame reads the caller's plane
                    assert (st.gbuffer.ptr, st.gbuffer.pitch_bytes) == (g.data_ptr(), w * 16), (f, i)
                else:
                    assert st.gbuffer.ptr != g.data_ptr(), (f, i)
                if keep and f > 0 and KEEP[f - 1]:
                    # frame f has completed: the G-buffer kept for it is no longer read
                    ins[i][f - 1]["gbuffer"].fill_(float("nan"))
        state = []
        for i, (w, h) in enumerate(SHAPES):
            st = ctx.stream_state(sids[i])
            hist = S.device_tensor_from_plane(st.history, h, w, 4, dtype)
            mom = S.device_tensor_from_plane(st.moments, h, w, ctx.moment_channels, dtype)
            state.append((_bits(hist), _bits(mom)))
    return [[_bits(o) for o in row] for row in outs], state


@pytest.mark.parametrize("storage,launch", [(1, 0), (1, 4), (0, 0)],
                         ids=["fp16_wide_per_stage", "fp16_wide_persistent", "fp32_per_stage"])
def test_kept_gbuffer_matches_saved_copy(S, storage, launch):
    dtype = torch.float16 if storage == S.STORAGE_FP16 else torch.float32
    copied, state_c = _run(S, storage, launch, dtype, keep=False)
    kept, state_k = _run(S, storage, launch, dtype, keep=True)
    for i in range(len(SHAPES)):
        for f in range(FRAMES):
            assert torch.equal(copied[i][f], kept[i][f]), (i, f, SHAPES[i])
        assert torch.equal(state_c[i][0], state_k[i][0]), ("history", i)
        assert torch.equal(state_c[i][1], state_k[i][1]), ("moments", i)


def test_request_flags_outside_the_defined_bits_are_rejected(S):
    dev = torch.device("cuda")
    w, h = 64, 48
    d = S.synth_frame(w, h, seed=5, stream=0, frame=0)
    item = {"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
            "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
            "motion": torch.from_numpy(d["motion"]).to(dev),
            "denoised": torch.empty((h, w, 4), dtype=torch.float16, device=dev)}
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=w, max_height=h, max_streams=1,
                   max_batch_frames=1, max_batch_pixels=w * h) as ctx:
        sid = ctx.stream_open(w, h)
        with pytest.raises(S.SvgfError, match="invalid argument"):
            ctx.denoise_batch([{"stream_id": sid, "flags": 0x2, **item}])
        ctx.denoise_batch([{"stream_id": sid, "flags": S.REQUEST_KEEP_GBUFFER, **item}])
        assert ctx.batch_status(ctx.batch_last(), wait=True) == 0
        torch.cuda.synchronize()
