"""Pins of the consumer-composite oracle (CPU only; oracle/composite_oracle.c).

The composite follows src/mesh.slang:1007 (history at the same pixel), :1161-1174 (reflection
hit reprojected through the previous view-projection, fallback hit_base * AMBIENT) and :1191
(final sum). The reference holds no test or vector for it, so the oracle is pinned by
construction: reflection hits are built by unprojecting chosen pixel centres through the
inverse camera in float64 (so a sign, axis, transposition or width/height mistake in the
oracle's projection lands on another pixel), by exact special cases (a hit on its own pixel
doubles the history; F = 0 leaves history + emissive; a behind-camera or off-screen hit falls
back to the ambient term), by the truncation of int2() at the frame's left and top edges, and
by numpy's float16 rounding for the half store.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import svgf_testutil as U

PLANES = ("history", "emissive", "fresnel", "reflection", "reflection_colour", "reflection_base")


def run(d, **kw):
    return O.composite(d["m"], *[d[k] for k in PLANES], **kw)


@pytest.mark.parametrize("w,h,seed,coherent", [(61, 37, 1, False), (160, 90, 2, False), (37, 61, 3, False),
                                                (8, 8, 4, False), (160, 90, 5, True), (64, 64, 6, True)])
def test_lookup_lands_on_the_constructed_pixel(w, h, seed, coherent):
    d = U.composite_scene(w, h, seed, coherent=coherent)
    refl = d["reflection"].reshape(-1, 4)
    tgt = d["target"].reshape(-1, 2)
    checked = 0
    for i in np.flatnonzero(refl[:, 3] == 2.0):
        got = O.composite_lookup(d["m"], w, h, refl[i, :3])
        want = None if tgt[i, 0] < 0 else (int(tgt[i, 0]), int(tgt[i, 1]))
        assert got == want, (i, got, want)
        checked += 1
    assert checked > 0.3 * w * h
    if coherent:  # floor hits land near the mirrored pixel
        tg = d["target"]
        ys, xs = np.nonzero(tg[..., 0] >= 0)
        assert np.abs(tg[ys, xs, 0] - xs).max() <= 2 and np.abs(tg[ys, xs, 1] - (h - 1 - ys)).max() <= 2


def test_lookup_is_not_symmetric_in_width_and_height():
    """A transposed frame must give different pixels: the construction is not symmetric."""
    d = U.composite_scene(160, 90, 2)
    refl = d["reflection"].reshape(-1, 4)
    tgt = d["target"].reshape(-1, 2)
    i = next(j for j in np.flatnonzero(refl[:, 3] == 2.0) if tgt[j, 0] > 100)
    assert O.composite_lookup(d["m"], 160, 90, refl[i, :3]) == (int(tgt[i, 0]), int(tgt[i, 1]))
    assert O.composite_lookup(d["m"], 90, 160, refl[i, :3]) != (int(tgt[i, 0]), int(tgt[i, 1]))


def test_lookup_truncation_and_edges_with_identity():
    """Identity matrix: clip = (x, y, z, 1), so u = x / 2 + 1/2, v = 1/2 - y / 2 exactly."""
    m = np.eye(4, dtype=np.float32)
    W, H = 8, 4
    # u W = -0.5 truncates to column 0 (inside); u W = -1 is off screen
    assert O.composite_lookup(m, W, H, (-1.125, 0.0, 0.0)) == (0, 2)
    assert O.composite_lookup(m, W, H, (-1.25, 0.0, 0.0)) is None
    # u W = 8 - 1/64 is column 7; u W = 8 is off screen
    assert O.composite_lookup(m, W, H, (1.0 - 1.0 / 256, 0.0, 0.0)) == (7, 2)
    assert O.composite_lookup(m, W, H, (1.0, 0.0, 0.0)) is None
    # rows count downward: y = +1 is the top row, y slightly above +1 truncates into row 0
    assert O.composite_lookup(m, W, H, (0.0, 1.0, 0.0)) == (4, 0)
    assert O.composite_lookup(m, W, H, (0.0, 1.25, 0.0)) == (4, 0)
    assert O.composite_lookup(m, W, H, (0.0, 1.5, 0.0)) is None
    assert O.composite_lookup(m, W, H, (0.0, -1.0, 0.0)) is None
    assert O.composite_lookup(m, W, H, (0.0, -0.999, 0.0)) == (4, 3)
    for bad in (np.nan, np.inf, -np.inf):
        assert O.composite_lookup(m, W, H, (bad, 0.0, 0.0)) is None
        assert O.composite_lookup(m, W, H, (0.0, bad, 0.0)) is None


def test_lookup_rejects_w_not_positive():
    m = np.eye(4, dtype=np.float32)
    m[3] = (0.0, 0.0, 1.0, 0.0)  # clip.w = z
    assert O.composite_lookup(m, 8, 8, (0.0, 0.0, 0.0)) is None    # w = 0
    assert O.composite_lookup(m, 8, 8, (0.0, 0.0, -2.0)) is None   # w < 0
    assert O.composite_lookup(m, 8, 8, (0.0, 0.0, 2.0)) == (4, 4)  # ndc (0, 0)
    # x / w: a point at x = 1, w = 2 is ndc 0.5 -> u = 0.75 -> column 6
    assert O.composite_lookup(m, 8, 8, (1.0, 0.0, 2.0)) == (6, 4)


def test_lookup_uses_the_translation_column():
    m = np.eye(4, dtype=np.float32)
    m[0, 3] = 0.5  # clip.x = x + 0.5
    assert O.composite_lookup(m, 8, 8, (0.0, 0.0, 0.0)) == (6, 4)
    m = np.eye(4, dtype=np.float32)
    m[3, 3] = 2.0  # clip.w = 2
    assert O.composite_lookup(m, 8, 8, (1.0, 0.0, 0.0)) == (6, 4)


def test_hit_on_own_pixel_doubles_the_history():
    w, h = 24, 16
    d = U.composite_scene(w, h, 7)
    M = d["m"].astype(np.float64).reshape(4, 4)
    ys, xs = np.mgrid[0:h, 0:w]
    pos = U.unproject(M, 2.0 * (xs + 0.5) / w - 1.0, 1.0 - 2.0 * (ys + 0.5) / h, np.full(xs.shape, 0.5))
    d["reflection"][..., :3] = pos.astype(np.float32)
    d["reflection"][..., 3] = 2.0
    d["emissive"][:] = 0.0
    d["fresnel"][:] = 1.0
    d["reflection_colour"][:] = 0.0
    out = run(d)
    assert np.array_equal(out[..., :3], 2.0 * d["history"][..., :3])
    assert np.all(out[..., 3] == 1.0)


def test_zero_fresnel_leaves_history_plus_emissive():
    d = U.composite_scene(40, 30, 8)
    d["fresnel"][:] = 0.0
    out = run(d)
    assert np.array_equal(out[..., :3], d["history"][..., :3] + d["emissive"][..., :3])


def test_kinds_and_fallback_terms():
    """Each pixel against the reference's sum with the lookup pixel taken from the scene
    construction, in float64 (one rounding per float op in the oracle: relative 1e-6)."""
    for w, h, seed in [(61, 37, 11), (160, 90, 12)]:
        d = U.composite_scene(w, h, seed)
        out = run(d)
        hist, emis, fres = (d[k][..., :3].astype(np.float64) for k in ("history", "emissive", "fresnel"))
        rcol, rbase = d["reflection_colour"][..., :3].astype(np.float64), d["reflection_base"][..., :3].astype(np.float64)
        kind = d["reflection"][..., 3]
        tgt = d["target"]
        refl = np.zeros_like(hist)
        ready = kind == 1.0
        refl[ready] = rcol[ready]
        rep = kind == 2.0
        hit = rep & (tgt[..., 0] >= 0)
        miss = rep & (tgt[..., 0] == -1)
        assert hit.sum() > 0 and miss.sum() > 0 and ready.sum() > 0
        refl[hit] = rcol[hit] + hist[tgt[hit][:, 1], tgt[hit][:, 0]]
        refl[miss] = rcol[miss] + rbase[miss] * np.float64(np.float32(0.03))
        want = hist + emis + refl * fres
        err = np.abs(out[..., :3] - want) / np.maximum(np.abs(want), 1e-6)
        assert err.max() < 1e-6, (w, h, float(err.max()))
        # no reflection (kind 0, NaN, 3, -1, 2.5, 1.5): exactly (history + emissive) + 0 * F
        none = ~(ready | rep)
        assert np.array_equal(out[none][:, :3], (d["history"] + d["emissive"])[none][:, :3])


def test_ambient_parameter_scales_the_fallback_only():
    d = U.composite_scene(61, 37, 13)
    a = run(d, ambient=0.03)
    b = run(d, ambient=0.5)
    miss = (d["reflection"][..., 3] == 2.0) & (d["target"][..., 0] == -1)
    assert np.array_equal(a[~miss], b[~miss])
    assert np.all(np.any(a[miss][:, :3] != b[miss][:, :3], axis=1) | (d["reflection_base"][miss][:, :3] == 0).all(1))


def test_half_store_is_numpy_float16_rounding():
    d = U.composite_scene(61, 37, 14, half=True)
    full = run(d)
    half = run(d, store=O.STORE_F16)
    assert np.array_equal(half, full.astype(np.float16).astype(np.float32))
