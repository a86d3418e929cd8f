"""GPU test of the frame profiler's hipEvent mode around the denoiser (the reference's GPU pass
profiler, src/main.cpp:2902-2953): log lines at the reference's frames, the average close to an
independent hipEvent timing of the same frames, and a wrapping pass at least as long as the pass
it wraps."""
from __future__ import annotations

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


def test_profiler_events_around_the_denoiser(S):
    w, h = 1280, 720
    d = S.synth_frame(w, h, seed=0, frame=1, pan_px=0.5)
    rad = torch.from_numpy(d["radiance"]).cuda().half()
    den = torch.empty((h, w, 4), dtype=torch.float16, device="cuda")
    mom = [torch.zeros((h, w, 2), dtype=torch.float16, device="cuda") for _ in range(2)]
    stream = torch.cuda.Stream()
    frames = 130
    logged = []
    with S.Context(S.MODE_REFERENCE, S.STORAGE_FP16, max_width=w, max_height=h) as ctx, \
            S.Profiler(["frame", "svgf"], device=0) as prof, torch.cuda.stream(stream):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k in range(frames):
            logged.append(prof.begin_frame())
            if k == 30:
                e0.record(stream)
            prof.pass_begin(0, stream)
            prof.pass_begin(1, stream)
            ctx.denoise(radiance=rad, denoised=den, moments_curr=mom[k % 2], moments_prev=mom[(k + 1) % 2])
            prof.pass_end(1, stream)
            prof.pass_end(0, stream)
            prof.end_frame()
        e1.record(stream)
        torch.cuda.synchronize()
        ema = prof.ema()
        line = prof.log_line()
    # frame counter starts at 16: lines at counters 60 and 120 = iterations 44 and 104
    assert [i for i, x in enumerate(logged) if x] == [44, 104]
    assert line.startswith("GPU times (ms): frame=") and " svgf=" in line
    per_frame_ms = e0.elapsed_time(e1) / (frames - 30)
    ema_ms = ema / 1e6
    print(f"profiler ema {ema_ms}, independent {per_frame_ms:.5f} ms/frame", flush=True)
    assert 0 < ema_ms[1] <= ema_ms[0] * 1.001
    # the events bracket only the denoise; the independent timing also holds the gaps between frames
    assert 0.3 * per_frame_ms < ema_ms[1] < 1.5 * per_frame_ms
