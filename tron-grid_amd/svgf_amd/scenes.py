"""Seeded synthetic inputs of the consumer composite (numpy, host only).

A composite frame is what TronGrid's mesh pass holds besides the denoised history
(src/mesh.slang:1007, 1161-1174, 1191): emissive, F_view, and per pixel a reflection whose
kind is reflection.w (2 = reproject the hit position through the previous view-projection,
1 = ready reflected colour, other = none). Reprojecting hits are built by unprojecting chosen
pixel centres through the inverse of a perspective camera in float64, so the pixel each one
must land on is known by construction.
"""
from __future__ import annotations

import numpy as np


def look_at_projection(w, h, eye, target):
    """Row-major float64 4x4 P V of a perspective camera at `eye` looking at `target` (60 degree
    vertical field of view, near 0.1, far 100, clip z in [0, w], clip w = -z_view)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    side = np.cross(fwd, [0.0, 1.0, 0.0])
    side /= np.linalg.norm(side)
    up = np.cross(side, fwd)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = side, up, -fwd
    V[:3, 3] = -V[:3, :3] @ eye
    t = 1.0 / np.tan(np.radians(30.0))
    n, f = 0.1, 100.0
    P = np.zeros((4, 4))
    P[0, 0], P[1, 1] = t * h / w, t
    P[2, 2], P[2, 3] = f / (n - f), f * n / (n - f)
    P[3, 2] = -1.0
    return P @ V


def view_projection(w, h, seed):
    """look_at_projection from a seeded position towards the origin; returns (M, eye)."""
    rng = np.random.default_rng(seed)
    eye = rng.uniform(-6.0, 6.0, 3) + np.array([0.0, 4.0, 0.0])
    target = rng.uniform(-1.0, 1.0, 3)
    return look_at_projection(w, h, eye, target), eye


def unproject(M, ndc_x, ndc_y, z_ndc):
    """World positions (..., 3) whose projection through M has the given NDC x, y and depth."""
    clip = np.stack([ndc_x, ndc_y, z_ndc, np.ones_like(ndc_x)], -1)
    wh = clip @ np.linalg.inv(M).T
    return wh[..., :3] / wh[..., 3:4]


def composite_scene(w, h, seed, half=False, coherent=False):
    """Seeded inputs of one composite frame. Returns a dict: m (16 float32, row-major), planes
    history, emissive, fresnel, reflection, reflection_colour, reflection_base ((h, w, 4)
    float32; the rgb planes hold half-representable values when half=True), and target ((h, w, 2)
    int: the history pixel each reprojecting reflection was built to land on, -1 for a miss,
    -2 for pixels that do not reproject).

    coherent=False (the stress mix) draws every pixel's reflection kind and target at random.
    coherent=True lays the frame out as a scene does: the top 30 % of rows show no reflection,
    the next 15 % a ready reflected colour (sky), the rest a reflective floor whose hits land
    near the pixel mirrored about the frame's middle row (a planar reflection in screen space,
    +-2 px of jitter), except the rightmost 5 % of columns (off screen) and 3 % of floor pixels
    (behind the camera)."""
    rng = np.random.default_rng(seed)
    M, eye = view_projection(w, h, seed)
    M32 = M.astype(np.float32)
    npx = w * h
    q = lambda a: a.astype(np.float16).astype(np.float32) if half else a.astype(np.float32)
    hist = q(rng.gamma(1.5, 2.0, (h, w, 4)))
    emis = np.zeros((h, w, 4), np.float32)
    lit = rng.random((h, w)) < 0.1
    emis[lit] = q(rng.uniform(0.0, 5.0, (int(lit.sum()), 4)))
    fres = q(rng.uniform(0.02, 1.0, (h, w, 4)))
    rcol = q(rng.uniform(0.0, 3.0, (h, w, 4)))
    rbase = q(rng.uniform(0.0, 1.0, (h, w, 4)))
    refl = np.zeros((h, w, 4), np.float32)
    refl[..., :3] = rng.uniform(-20.0, 20.0, (h, w, 3))
    target = np.full((h, w, 2), -2, np.int64)

    u = rng.random(npx)
    kinds = np.zeros(npx)
    if coherent:
        ys, xs = np.divmod(np.arange(npx), w)
        cls = np.full(npx, 7)
        cls[ys >= 0.3 * h] = 5
        floor = ys >= 0.45 * h
        cls[floor] = 0
        cls[floor & (xs >= 0.95 * w)] = 2
        cls[floor & (xs < 0.95 * w) & (u < 0.03)] = 1
        tx = np.clip(xs + rng.integers(-2, 3, npx), 0, w - 1)
        ty = np.clip(h - 1 - ys + rng.integers(-2, 3, npx), 0, h - 1)
    else:
        # 40 % reproject into the frame, 5 % behind the camera, 5 % off screen, 3 % into the
        # truncation band left of / above the frame, 2 % just inside the right / bottom edge,
        # 15 % ready colour, 3 % other kind values (no reflection), the rest no reflection
        cuts = np.cumsum([0.40, 0.05, 0.05, 0.03, 0.02, 0.15, 0.03])
        cls = np.searchsorted(cuts, u, side="right")
        tx = rng.integers(0, w, npx)
        ty = rng.integers(0, h, npx)
    zn = rng.uniform(0.05, 0.95, npx)
    fx = tx + 0.5
    fy = ty + 0.5
    fx = np.where(cls == 3, -0.5, fx)
    fy = np.where((cls == 3) & (rng.random(npx) < 0.5), -0.5, fy)
    fx = np.where(cls == 4, w - 0.25, fx)
    fy = np.where((cls == 4) & (rng.random(npx) < 0.5), h - 0.25, fy)
    ndc_x = 2.0 * fx / w - 1.0
    ndc_y = 1.0 - 2.0 * fy / h
    ndc_x = np.where(cls == 2, np.sign(rng.random(npx) - 0.5) * rng.uniform(1.1, 3.0, npx), ndc_x)
    pos = unproject(M, ndc_x, ndc_y, zn)
    behind = cls == 1
    pos[behind] = eye - (pos[behind] - eye) * rng.uniform(0.5, 2.0, (int(behind.sum()), 1))
    pos32 = pos.astype(np.float32)
    # the float64 projection of the float32 position decides the expected pixel
    clip = np.concatenate([pos32.astype(np.float64), np.ones((npx, 1))], 1) @ M32.astype(np.float64).T
    with np.errstate(all="ignore"):
        px = np.floor(np.where(clip[:, 3] > 0, (clip[:, 0] / clip[:, 3] * 0.5 + 0.5) * w, -9)).astype(np.int64)
        py = np.floor(np.where(clip[:, 3] > 0, (clip[:, 1] / clip[:, 3] * -0.5 + 0.5) * h, -9)).astype(np.int64)
    # trunc toward zero: the band (-1, 0) maps to 0
    px = np.where((px == -1) & (cls == 3), 0, px)
    py = np.where((py == -1) & (cls == 3), 0, py)
    inside = (clip[:, 3] > 0) & (px >= 0) & (px < w) & (py >= 0) & (py < h)
    reproj = cls <= 4
    kinds[reproj] = 2.0
    kinds[cls == 5] = 1.0
    kinds[cls == 6] = np.array([3.0, -1.0, np.nan, 2.5, 1.5])[rng.integers(0, 5, int((cls == 6).sum()))]
    flat = refl.reshape(-1, 4)
    flat[reproj, :3] = pos32[reproj]
    flat[:, 3] = kinds
    tflat = target.reshape(-1, 2)
    tflat[reproj] = np.where(inside[reproj, None], np.stack([px, py], 1)[reproj], -1)
    return {"m": M32.reshape(16), "history": hist, "emissive": emis, "fresnel": fres, "reflection": refl,
            "reflection_colour": rcol, "reflection_base": rbase, "target": target}


def pack_octahedral(n):
    """The reference's 16 + 16-bit octahedral code of unit vectors n (..., 3)
    (src/mesh.slang:327-339), evaluated in float64: the producer's encoder."""
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    s = np.abs(n).sum(-1, keepdims=True)
    p = n[..., :2] / np.maximum(s, 1e-6)
    snap = np.where(p >= 0.0, 1.0, -1.0)
    fold = (1.0 - np.abs(p[..., ::-1])) * snap
    p = np.where(n[..., 2:3] < 0.0, fold, p)
    q = np.floor(np.clip(p * 0.5 + 0.5, 0.0, 1.0) * 65535.0 + 0.5).astype(np.uint32)
    return q[..., 0] | (q[..., 1] << 16)


def pack_scene(w, h, seed, half=False, move=0.02):
    """Seeded producer inputs of one frame: a camera (vp) that moved by `move` metres since the
    previous frame (pvp), world positions of the pixel centres at random depths with coverage in
    position.w (the top quarter of the rows is sky, 10 % of the others transparent-only, both
    uncovered; 2 % of covered pixels sit behind the previous camera), octahedral-coded normals,
    and the lighting terms indirect, direct_diffuse (ao in .w), direct_specular ((h, w, 4);
    half-representable when half=True). Also returns depth ((h, w) float64: clip.w of the
    float32 position through the float32 vp) and normal ((h, w, 3) float64 unit normals)."""
    rng = np.random.default_rng(seed)
    eye = rng.uniform(-6.0, 6.0, 3) + np.array([0.0, 4.0, 0.0])
    target = rng.uniform(-1.0, 1.0, 3)
    vp = look_at_projection(w, h, eye, target)
    d = rng.normal(size=3)
    peye = eye + move * d / np.linalg.norm(d)
    pvp = look_at_projection(w, h, peye, target + 0.5 * move * rng.normal(size=3))
    vp32, pvp32 = vp.astype(np.float32), pvp.astype(np.float32)
    ys, xs = np.mgrid[0:h, 0:w]
    zn = rng.uniform(0.02, 0.98, (h, w))
    pos = np.zeros((h, w, 4), np.float32)
    world = unproject(vp, 2.0 * (xs + 0.5) / w - 1.0, 1.0 - 2.0 * (ys + 0.5) / h, zn)
    covered = (ys >= h // 4) & (rng.random((h, w)) >= 0.1)
    behind = covered & (rng.random((h, w)) < 0.02)
    fwd_prev = target - peye
    fwd_prev /= np.linalg.norm(fwd_prev)
    world[behind] = peye - fwd_prev * rng.uniform(0.5, 5.0, (int(behind.sum()), 1))
    pos[..., :3] = world
    pos[..., 3] = covered.astype(np.float32)
    pos[~covered, :3] = rng.uniform(-50.0, 50.0, (int((~covered).sum()), 3))
    normal = rng.normal(size=(h, w, 3))
    normal /= np.linalg.norm(normal, axis=-1, keepdims=True)
    q = (lambda a: a.astype(np.float16).astype(np.float32)) if half else (lambda a: a.astype(np.float32))
    ind = q(rng.gamma(1.0, 1.0, (h, w, 4)))
    dd = q(rng.gamma(1.0, 0.5, (h, w, 4)))
    dd[..., 3] = q(rng.uniform(0.0, 1.0, (h, w)))
    ds = q(rng.gamma(0.5, 0.2, (h, w, 4)))
    p64 = np.concatenate([pos[..., :3].astype(np.float64), np.ones((h, w, 1))], -1)
    depth = p64 @ vp32.astype(np.float64)[3]
    return {"vp": vp32.reshape(16), "pvp": pvp32.reshape(16), "position": pos, "normal_oct": pack_octahedral(normal),
            "indirect": ind, "direct_diffuse": dd, "direct_specular": ds, "depth": depth, "normal": normal,
            "covered": covered, "behind": behind}
