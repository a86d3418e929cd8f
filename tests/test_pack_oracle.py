"""Pins of the producer-pack oracle (CPU only; oracle/pack_oracle.c).

The pack writes the denoiser input where opaque geometry covers a pixel and zero elsewhere
(src/main.cpp:595-646, src/mesh.slang:995-998), unpacks the reference's octahedral normal
(mesh.slang:341-352), and derives linear depth and motion from the world position through the
current and previous view-projections (mesh.slang:589-591). The reference ships no vector for
any of it; the oracle is pinned by the reference's own encoder (mesh.slang:327-339, restated in
float64) round-tripping through the oracle's decoder, by positions unprojected from pixel
centres (a static camera has zero motion, a pure NDC shift a known one, the depth is the
constructed view distance), by the exact zero of uncovered pixels, and by numpy's float16
rounding for the half store.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from svgf_amd.scenes import pack_octahedral, pack_scene

LIGHT = ("indirect", "direct_diffuse", "direct_specular")


def run(d, vp=None, pvp=None, **kw):
    return O.pack(d["vp"] if vp is None else vp, d["pvp"] if pvp is None else pvp, d["position"], d["normal_oct"],
                  *[d[k] for k in LIGHT], **kw)


def test_octahedral_round_trip_of_the_reference_encoder():
    rng = np.random.default_rng(3)
    n = rng.normal(size=(20000, 3))
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                     [1, 1, -1e-3], [-1, 1, -0.5], [0.3, -0.2, -0.9]], np.float64)
    n = np.concatenate([axes, n])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    codes = pack_octahedral(n)
    got = np.stack([O.unpack_octahedral(c) for c in codes])
    # 16-bit quantisation of the octahedral square: well below 1e-4 per component
    assert np.abs(got - n).max() < 1e-4
    assert np.abs(np.linalg.norm(got.astype(np.float64), axis=1) - 1.0).max() < 3e-7
    # the lower hemisphere folds (n.z < 0) land there too
    assert np.all(np.sign(got[n[:, 2] < -1e-3, 2]) < 0)


def test_depth_is_the_view_distance():
    d = pack_scene(96, 64, 1)
    _, g, _ = run(d)
    cov = d["covered"] & ~d["behind"]
    rel = np.abs(g[..., 3][cov] - d["depth"][cov]) / d["depth"][cov]
    assert rel.max() < 2e-5
    assert np.all(g[..., 3][cov] > 0.09)  # in front of the near plane
    nrm = np.abs(g[..., :3][d["covered"]] - d["normal"][d["covered"]])
    assert nrm.max() < 1e-4


def test_static_camera_has_zero_motion():
    d = pack_scene(120, 80, 2)
    _, _, mv = run(d, pvp=d["vp"])
    cov = d["covered"] & ~d["behind"]
    assert np.abs(mv[cov]).max() < 2e-3


@pytest.mark.parametrize("dx,dy", [(0.25, 0.0), (0.0, -0.125), (-0.1, 0.3)])
def test_ndc_shift_gives_the_pixel_shift(dx, dy):
    """pvp = T vp with T adding (dx, dy) times clip.w to clip.xy: the previous NDC is shifted by
    (dx, dy), i.e. by (dx W / 2, -dy H / 2) pixels (rows grow downward)."""
    w, h = 128, 72
    d = pack_scene(w, h, 4)
    T = np.eye(4)
    T[0, 3], T[1, 3] = dx, dy
    pvp = (T @ d["vp"].astype(np.float64).reshape(4, 4)).astype(np.float32)
    _, _, mv = run(d, pvp=pvp)
    cov = d["covered"] & ~d["behind"]
    assert np.abs(mv[cov][:, 0] - dx * w / 2).max() < 3e-3
    assert np.abs(mv[cov][:, 1] + dy * h / 2).max() < 3e-3


def test_uncovered_and_behind_pixels():
    d = pack_scene(64, 48, 5)
    rad, g, mv = run(d)
    unc = ~d["covered"]
    assert unc.sum() > 0 and d["behind"].sum() > 0
    assert np.all(rad[unc] == 0.0) and np.all(g[unc] == 0.0)
    assert np.all(np.signbit(rad[unc]) == 0)  # +0, as the clear writes
    assert np.all(mv[unc] == O.PACK_NO_MOTION) and np.all(mv[d["behind"]] == O.PACK_NO_MOTION)
    # coverage is position.w > 0: NaN and negative weights are uncovered
    d2 = dict(d)
    d2["position"] = d["position"].copy()
    d2["position"][0, 0, 3] = np.nan
    d2["position"][0, 1, 3] = -1.0
    d2["position"][0, 2, 3] = 0.5
    r2, _, _ = run(d2)
    assert np.all(r2[0, 0] == 0) and np.all(r2[0, 1] == 0) and r2[0, 2, 3] == 1.0


def test_radiance_is_the_mesh_pass_sum():
    d = pack_scene(80, 60, 6)
    rad, _, _ = run(d)
    cov = d["covered"]
    ind, dd, ds = (d[k].astype(np.float64) for k in LIGHT)
    want = (ind[..., :3] + dd[..., :3]) * dd[..., 3:4] + ds[..., :3]
    rel = np.abs(rad[cov][:, :3] - want[cov]) / np.maximum(np.abs(want[cov]), 1e-6)
    assert rel.max() < 5e-7
    assert np.all(rad[cov][:, 3] == 1.0)


def test_runaway_motion_is_clamped_to_no_motion():
    d = pack_scene(32, 24, 7)
    pvp = d["vp"].reshape(4, 4).copy()
    pvp[0] *= 1e8  # clip.x huge: the previous position runs off beyond 1e6 px
    pvp[1] *= 1e8
    _, _, mv = run(d, pvp=pvp.reshape(16))
    cov = d["covered"] & ~d["behind"]
    assert np.all(mv[cov] == O.PACK_NO_MOTION)


def test_half_store_is_numpy_float16_rounding():
    d = pack_scene(61, 37, 8, half=True)
    full = run(d)[0]
    half = run(d, store=O.STORE_F16)[0]
    assert np.array_equal(half, full.astype(np.float16).astype(np.float32))
