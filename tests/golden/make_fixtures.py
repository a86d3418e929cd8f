"""Writes the golden fixtures of this directory from the FP64 restatement in svgf_restate.py.

    python tests/golden/make_fixtures.py

Inputs are seeded numpy scenes, independent of the library's synthetic frames. Radiance is
rounded to binary16 so that the FP16 storage forms read exactly the same inputs.

reference.npz
    REFERENCE mode, 3 a-trous passes, over 3 frames per shape: a cold first frame, then warm
    frames that read the previous frame's moments.
canonical.npz
    CANONICAL mode with the default parameters, 5 iterations (strides up to 16 on frames
    smaller than that), over 5 frames per shape. Each scene has a depth step, normals that vary
    enough for the n.n'^128 term to matter, a fractional pan, and pixels that are disoccluded
    (depth jump), that reproject out of the frame, or that move by whole pixels. Each frame
    gets the raw-moment and the compensated-moment outputs.

Per shape i, arrays are named s{i}_<plane>, stacked over frames.
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import svgf_restate as R  # noqa: E402

REFERENCE_SHAPES = [(8, 6), (13, 7), (1, 1), (2, 9)]  # (w, h)
REFERENCE_FRAMES = 3
REFERENCE_PASSES = 3
CANONICAL_SHAPES = [(8, 6), (13, 7), (2, 9), (17, 11), (1, 1)]
CANONICAL_FRAMES = 5


def _half(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def reference_inputs(w, h, seed):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(REFERENCE_FRAMES):
        rad = np.ones((h, w, 4), np.float32)
        grad = 1.0 + 0.4 * np.arange(w)[None, :] + 0.2 * np.arange(h)[:, None]
        rad[..., :3] = _half(grad[..., None] * rng.exponential(1.0, (h, w, 3)))
        out.append(rad)
    return out


def canonical_inputs(w, h, seed):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    n = np.stack([0.05 * np.sin(0.7 * X + 0.3 * Y), 0.04 * np.cos(0.5 * Y - 0.2 * X), np.ones_like(X)], -1)
    n = (n / np.linalg.norm(n, axis=-1, keepdims=True)).astype(np.float32)
    z0 = 5.0 + 0.05 * X + 0.03 * Y
    z0 = np.where(X >= w // 2, z0 + 4.0, z0)  # depth step
    frames = []
    for f in range(CANONICAL_FRAMES):
        g = np.empty((h, w, 4), np.float32)
        g[..., :3] = n
        z = z0 * (1.0 + 0.01 * np.sin(f + 0.1 * X))
        jump = rng.random((h, w)) < 0.12  # disoccluded this frame
        g[..., 3] = np.where(jump, z * 1.5, z).astype(np.float32)
        mv = np.empty((h, w, 2), np.float32)
        mv[..., 0], mv[..., 1] = 0.37, -0.61
        r = rng.random((h, w))
        mv[r < 0.08] = (-40.0, 0.0)  # reprojects out of the frame
        mv[(r >= 0.08) & (r < 0.16)] = (1.0, 0.0)  # whole-pixel motion
        rad = np.ones((h, w, 4), np.float32)
        grad = 1.0 + 0.3 * X + 0.1 * Y
        rad[..., :3] = _half(grad[..., None] * rng.exponential(1.0, (h, w, 3)))
        if w * h > 4:
            rad[h // 2, w // 3, :3] = 40.0  # a firefly
        frames.append({"radiance": rad, "gbuffer": g, "motion": mv})
    return frames


def reference_fixture():
    data = {"shapes": np.array(REFERENCE_SHAPES, np.int32), "passes": np.int32(REFERENCE_PASSES)}
    for i, (w, h) in enumerate(REFERENCE_SHAPES):
        ins = reference_inputs(w, h, 100 + i)
        mom, dens, moms = None, [], []
        for rad in ins:
            mom, den = R.reference_frame(rad, mom, REFERENCE_PASSES)
            moms.append(mom)
            dens.append(den)
        data[f"s{i}_radiance"] = np.stack(ins)
        data[f"s{i}_moments"] = np.stack(moms)
        data[f"s{i}_denoised"] = np.stack(dens)
    return data


def canonical_fixture():
    data = {"shapes": np.array(CANONICAL_SHAPES, np.int32)}
    for i, (w, h) in enumerate(CANONICAL_SHAPES):
        ins = canonical_inputs(w, h, 200 + i)
        for k in ("radiance", "gbuffer", "motion"):
            data[f"s{i}_{k}"] = np.stack([d[k] for d in ins])
        for form, comp in (("raw", False), ("comp", True)):
            hist = mom = gp = None
            hs, ms, ds = [], [], []
            for f, d in enumerate(ins):
                hist, mom, den = R.canonical_frame(d["radiance"], d["gbuffer"], gp, d["motion"], hist, mom, comp,
                                                   history_valid=f > 0)
                gp = d["gbuffer"]
                hs.append(hist)
                ms.append(mom)
                ds.append(den)
            data[f"s{i}_{form}_history"] = np.stack(hs)
            data[f"s{i}_{form}_moments"] = np.stack(ms)
            data[f"s{i}_{form}_denoised"] = np.stack(ds)
    return data


def main():
    np.savez_compressed(HERE / "reference.npz", **reference_fixture())
    np.savez_compressed(HERE / "canonical.npz", **canonical_fixture())
    for n in ("reference.npz", "canonical.npz"):
        print(n, (HERE / n).stat().st_size, "bytes")


if __name__ == "__main__":
    main()
