"""Shared helpers of the test suite: error metrics and sequence drivers (GPU via the C ABI,
CPU via the oracle)."""
from __future__ import annotations

import numpy as np

REL_FLOOR = 1e-3  # cd/m^2: relative errors are |a - b| / max(|b|, REL_FLOOR)


def rel_err(a, b, floor=REL_FLOOR):
    """(max, mean) of |a - b| / max(|b|, floor) over every element."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    r = np.abs(a - b) / np.maximum(np.abs(b), floor)
    return float(r.max()), float(r.mean())


def seq_inputs(w, h, frames, *, seed0=0, stream=0, noise=0, pan_px=0.5, disocclusion=0.0, half=False):
    """Synthetic inputs for `frames` consecutive frames (frame f uses seed seed0 + f)."""
    import svgf_amd as S

    out = []
    for f in range(frames):
        d = S.synth_frame(w, h, seed=seed0 + f, stream=stream, frame=f, noise=noise, pan_px=pan_px,
                          disocclusion=disocclusion)
        if half:
            d["radiance"] = d["radiance"].astype(np.float16).astype(np.float32)
        out.append(d)
    return out


def oracle_canonical_sequence(inputs, params, threads=1):
    """Run the FP64 canonical oracle over a sequence; returns per-frame output dicts."""
    import oracle as O

    hist = mom = None
    gprev = None
    outs = []
    for f, d in enumerate(inputs):
        o = O.c_frame_f64(d["radiance"], d["gbuffer"], gprev, d["motion"], hist, mom, params,
                          history_valid=f > 0, debug=True, threads=threads)
        hist, mom, gprev = o["history"], o["moments"], d["gbuffer"]
        outs.append(o)
    return outs


def decode_state(ctx, hist, mom):
    """History (h, w, 4: colour, length) and moments (h, w, 2: the form's two values) of a
    context's state planes as float64 numpy arrays (the wide state decoded)."""
    import svgf_amd as S

    hn = hist.cpu().numpy()
    mn = mom.cpu().numpy()
    if not ctx.wide:
        return hn.astype(np.float64), mn.astype(np.float64)
    rgb, n = S.decode_history(hn)
    return np.concatenate([rgb, n[..., None].astype(np.float64)], axis=-1), S.decode_moments(mn)


def moment_errors(got, ref):
    """Errors of compensated moments (mean, sigma) against the oracle's: the mean and the raw
    second moment E[x^2] = mean^2 + sigma^2 (what the reference stores), relative with floors of
    1e-3 cd/m^2 and its square. (Sigma alone is ill-conditioned where sigma << |mean|.)"""
    m2g = got[..., 0] ** 2 + got[..., 1] ** 2
    m2r = ref[..., 0] ** 2 + ref[..., 1] ** 2
    return {"mean": rel_err(got[..., 0], ref[..., 0]), "second": rel_err(m2g, m2r, REL_FLOOR ** 2)}


def gpu_canonical_sequence(ctx, inputs, dtype, crops=None):
    """Run the GPU canonical form (single-frame C ABI) over a sequence with caller-owned
    ping-pong planes; returns per-frame numpy dicts (denoised, history, moments; the wide FP16
    state decoded: history = (colour, length), moments = the form's two values). With `crops`,
    a list of (y0, x0, c), each frame gives a list of such dicts, one per c x c crop."""
    import torch

    dev = torch.device("cuda")
    h, w = inputs[0]["radiance"].shape[:2]
    hist = [torch.zeros((h, w, 4), dtype=dtype, device=dev) for _ in range(2)]
    mom = [torch.zeros((h, w, ctx.moment_channels), dtype=dtype, device=dev) for _ in range(2)]
    gb = [None, None]
    outs = []
    for f, d in enumerate(inputs):
        cur, prev = f % 2, (f + 1) % 2
        rad = torch.from_numpy(d["radiance"]).to(dev).to(dtype)
        gb[cur] = torch.from_numpy(d["gbuffer"]).to(dev)
        mv = torch.from_numpy(d["motion"]).to(dev)
        den = torch.empty((h, w, 4), dtype=dtype, device=dev)
        ctx.denoise(radiance=rad, gbuffer=gb[cur], gbuffer_prev=gb[prev] if f else None, motion=mv,
                    history_prev=hist[prev] if f else None, history_curr=hist[cur],
                    moments_prev=mom[prev] if f else None, moments_curr=mom[cur], denoised=den,
                    history_valid=f > 0)
        torch.cuda.synchronize()
        if crops is None:
            hd, md = decode_state(ctx, hist[cur], mom[cur])
            outs.append({"denoised": den.float().cpu().numpy(), "history": hd, "moments": md})
            continue
        per = []
        for y0, x0, c in crops:
            win = (slice(y0, y0 + c), slice(x0, x0 + c))
            hd, md = decode_state(ctx, hist[cur][win], mom[cur][win])
            per.append({"denoised": den[win].float().cpu().numpy(), "history": hd, "moments": md})
        outs.append(per)
    return outs


# Receptive field of a CANONICAL output after 3 frames: 5 a-trous passes reach 2s each
# (taps and the variance prefilter are both centred on the pixel; s = 1..16) = 62, the depth
# gradient and the spatial variance 1 each, the reprojection 2, and each earlier frame's
# history (reprojection 2 + variance 1 + first pass 2) 5 more: under 80 in all. On the
# 640x520 oracle sequence of the locality test the farthest pixel a crop edge changes is 68.
MARGIN = 128
# The same after 16 frames at the 4K precision sequence's pan (3.74 px/frame at 3840x2160), lognormal noise and
# 10% disocclusion: the farthest pixel a crop edge changes is 127 (denoised), 78 (history),
# 59 (moments) on a 1100x700 oracle sequence; 192 leaves a third more.
MARGIN_16 = 192


def crop_windows(w, h, c):
    """(y0, x0) of c x c crops at the top-left corner, the centre and the bottom-right corner."""
    return [(0, 0), ((h - c) // 2, (w - c) // 2), (h - c, w - c)]


def crop_interior(y0, x0, c, w, h, margin=MARGIN):
    """Slices of a crop that are beyond the receptive field of every crop edge that is not an
    image edge (clamping and reprojection rejection agree there)."""
    top = 0 if y0 == 0 else margin
    left = 0 if x0 == 0 else margin
    bottom = c if y0 + c == h else c - margin
    right = c if x0 + c == w else c - margin
    return slice(top, bottom), slice(left, right)


# ---------------------------------------------------------------------------------------------
# consumer composite scenes (the generator lives with the library's other synthetic inputs)
# ---------------------------------------------------------------------------------------------
from svgf_amd.scenes import composite_scene, unproject, view_projection  # noqa: E402,F401
