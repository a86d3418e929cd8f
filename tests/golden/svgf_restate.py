"""Independent FP64 restatement of both SVGF modes, used to write the golden fixtures.

Test infrastructure only. It is written from the semantics of the pass, not from the C oracle:
whole-image numpy arrays with clamped index grids instead of per-pixel loops.

REFERENCE mode follows svgf_denoise.slang:
  - pass 0 (:176-191): same-pixel moment EMA, and sigma = sqrt(max(var, 0) + eps) in alpha;
  - passes 1..N (:209-274): 5x5 B3 taps at stride 2^(pass-1), with luminance edge stop
    exp(-|l_q - l_p| / max(phi sigma_p, eps)), clamp-to-edge taps and the centre fallback when
    the weight sum is <= 1e-6. Alpha carries sigma, and the last pass writes alpha = 1.

CANONICAL mode is the SVGF of Schied et al. 2017 with this build's choices:
  - bilinear reprojection with per-tap depth / normal / history-length rejection;
  - colour and moment EMAs with alpha = max(alpha, 1/n);
  - raw moments (mean, E[l^2]) or compensated ones (mean, sqrt of the centred second moment);
  - the 3x3 edge-stopped spatial variance for pixels with n < 4;
  - N a-trous iterations whose variance prefilter reads the lattice neighbours at +-s, with
    sigma_l = phi_l sqrt(g + eps), depth and normal edge stops, and the variance filtered by
    the squared weights.
"""
from __future__ import annotations

import numpy as np

# Rec.709 luminance and the B3 kernel, as the shader's float constants
LUMA = np.array([np.float32(0.2126), np.float32(0.7152), np.float32(0.0722)], np.float64)
B3 = np.array([1, 4, 6, 4, 1], np.float64) / 16.0
BIN3 = np.array([1, 2, 1], np.float64) / 4.0


def f32(v):
    """A parameter as the float32 the pass receives, widened to float64."""
    return float(np.float32(v))


def luminance(img):
    return img[..., 0] * LUMA[0] + img[..., 1] * LUMA[1] + img[..., 2] * LUMA[2]


def shifted(img, dx, dy):
    """img sampled at (x + dx, y + dy), coordinates clamped into the image."""
    h, w = img.shape[:2]
    ys = np.clip(np.arange(h) + dy, 0, h - 1)
    xs = np.clip(np.arange(w) + dx, 0, w - 1)
    return img[ys][:, xs]


# ------------------------------------------------------------------------------------------
# REFERENCE mode
# ------------------------------------------------------------------------------------------
def reference_frame(radiance, moments_prev, iterations, phi=2.0, alpha=0.2, eps=1e-4):
    """One frame of svgf_denoise.slang: returns (moments_curr (h,w,2), denoised (h,w,4))."""
    alpha, eps, phi = f32(alpha), f32(eps), f32(phi)
    rad = np.asarray(radiance, np.float64)
    h, w = rad.shape[:2]
    m = np.zeros((h, w, 2)) if moments_prev is None else np.asarray(moments_prev, np.float64)
    l = luminance(rad)
    mu1 = m[..., 0] + alpha * (l - m[..., 0])
    mu2 = m[..., 1] + alpha * (l * l - m[..., 1])
    cur = np.empty((h, w, 4))
    cur[..., :3] = rad[..., :3]
    cur[..., 3] = np.sqrt(np.maximum(mu2 - mu1 * mu1, 0.0) + eps)
    for p in range(1, iterations + 1):
        s = 1 << (p - 1)
        lc = luminance(cur)
        scale = np.maximum(phi * cur[..., 3], eps)
        acc = np.zeros((h, w, 3))
        wsum = np.zeros((h, w))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                q = shifted(cur, dx * s, dy * s)
                wt = B3[dx + 2] * B3[dy + 2] * np.exp(-np.abs(luminance(q) - lc) / scale)
                acc += wt[..., None] * q[..., :3]
                wsum += wt
        nxt = np.empty_like(cur)
        ok = wsum > 1e-6
        with np.errstate(invalid="ignore", divide="ignore"):
            nxt[..., :3] = np.where(ok[..., None], acc / wsum[..., None], cur[..., :3])
        nxt[..., 3] = 1.0 if p == iterations else cur[..., 3]
        cur = nxt
    return np.stack([mu1, mu2], -1), cur


# ------------------------------------------------------------------------------------------
# CANONICAL mode
# ------------------------------------------------------------------------------------------
CANONICAL_DEFAULTS = dict(iterations=5, phi_luma=4.0, phi_normal=128.0, phi_depth=1.0, alpha_color=0.2,
                          alpha_moments=0.2, variance_eps=1e-4, depth_eps=1e-3, depth_reject=0.1,
                          normal_reject=0.9, history_cap=32)


def _edge_terms(P, g, gshift, ox, oy, gx, gy):
    """(depth argument, normal factor) of the tap at offset (ox, oy) for every pixel."""
    arg = np.zeros(g.shape[:2])
    if P["phi_depth"] > 0.0:
        den = P["phi_depth"] * np.abs(gx * ox + gy * oy) + P["depth_eps"]
        arg = np.abs(gshift[..., 3].astype(np.float64) - g[..., 3]) / den
    fac = np.ones(g.shape[:2])
    if P["phi_normal"] > 0.0:
        d = np.einsum("hwc,hwc->hw", g[..., :3].astype(np.float64), gshift[..., :3].astype(np.float64))
        fac = np.where(d > 0.0, np.power(np.maximum(d, 1e-300), P["phi_normal"]), 0.0)
    return arg, fac


def _depth_gradient(g):
    z = g[..., 3].astype(np.float64)
    gx = 0.5 * (shifted(z, 1, 0) - shifted(z, -1, 0))
    gy = 0.5 * (shifted(z, 0, 1) - shifted(z, 0, -1))
    return gx, gy


def canonical_frame(radiance, gbuffer, gbuffer_prev, motion, history_prev, moments_prev, compensated,
                    history_valid, **params):
    """One CANONICAL frame. history_prev (h,w,4): colour + length; moments_prev (h,w,2): raw
    (mean, E[l^2]) or compensated (mean, sigma). Returns history_curr, moments_curr, denoised."""
    P = {k: (f32(v) if isinstance(v, float) else v) for k, v in {**CANONICAL_DEFAULTS, **params}.items()}
    rad = np.asarray(radiance, np.float64)
    g = np.asarray(gbuffer, np.float32)
    h, w = rad.shape[:2]
    l = luminance(rad)
    z = g[..., 3].astype(np.float64)

    # ---- temporal: bilinear taps of the reprojected position, each tested on its own ----
    sumb = np.zeros((h, w))
    hc = np.zeros((h, w, 3))
    m_a = np.zeros((h, w))
    m_b = np.zeros((h, w))
    taps = []
    nmax = np.zeros((h, w), np.int64)
    if history_valid:
        gp = np.asarray(gbuffer_prev, np.float32)
        hp = np.asarray(history_prev, np.float64)
        mp = np.asarray(moments_prev, np.float64)
        mv = np.asarray(motion, np.float32).astype(np.float64)
        fl = np.floor(mv)
        frac = mv - fl
        X, Y = np.meshgrid(np.arange(w), np.arange(h))
        x0 = X + fl[..., 0].astype(np.int64)
        y0 = Y + fl[..., 1].astype(np.int64)
        for jj in (0, 1):
            for ii in (0, 1):
                b = (frac[..., 0] if ii else 1.0 - frac[..., 0]) * (frac[..., 1] if jj else 1.0 - frac[..., 1])
                qx, qy = x0 + ii, y0 + jj
                inside = (qx >= 0) & (qx < w) & (qy >= 0) & (qy < h)
                cx, cy = np.clip(qx, 0, w - 1), np.clip(qy, 0, h - 1)
                gq = gp[cy, cx]
                dz = np.abs(gq[..., 3].astype(np.float64) - z)
                nd = np.einsum("hwc,hwc->hw", gq[..., :3].astype(np.float64), g[..., :3].astype(np.float64))
                nq = hp[cy, cx, 3]
                use = (b > 0.0) & inside & (dz <= P["depth_reject"] * z) & (nd >= P["normal_reject"]) & (nq >= 1)
                bu = np.where(use, b, 0.0)
                sumb += bu
                hc += bu[..., None] * hp[cy, cx, :3]
                nmax = np.maximum(nmax, np.where(use, nq, 0).astype(np.int64))
                taps.append((bu, mp[cy, cx, 0], mp[cy, cx, 1]))
    hist = sumb > 0.0
    safe = np.where(hist, sumb, 1.0)
    hc = np.where(hist[..., None], hc / safe[..., None], 0.0)
    n_new = np.where(hist, np.minimum(nmax + 1, P["history_cap"]), 1)
    ac = np.maximum(P["alpha_color"], 1.0 / n_new)
    am = np.maximum(P["alpha_moments"], 1.0 / n_new)
    colour = hc + ac[..., None] * (rad[..., :3] - hc)
    if compensated:
        mu = sum(bu * ma for bu, ma, _ in taps) / safe if taps else np.zeros((h, w))
        m2 = sum(bu * (mb * mb + (ma - mu) ** 2) for bu, ma, mb in taps) / safe if taps else np.zeros((h, w))
        mu = np.where(hist, mu, 0.0)
        m2 = np.where(hist, m2, 0.0)
        d = l - mu
        var = (1.0 - am) * (m2 + am * d * d)
        mom = np.stack([mu + am * d, np.sqrt(var)], -1)
    else:
        h1 = sum(bu * ma for bu, ma, _ in taps) / safe if taps else np.zeros((h, w))
        h2 = sum(bu * mb for bu, _, mb in taps) / safe if taps else np.zeros((h, w))
        h1 = np.where(hist, h1, 0.0)
        h2 = np.where(hist, h2, 0.0)
        mu1 = h1 + am * (l - h1)
        mu2 = h2 + am * (l * l - h2)
        mom = np.stack([mu1, mu2], -1)
        var = np.maximum(mu2 - mu1 * mu1, 0.0)

    # ---- spatial variance of young pixels: 3x3 window of temporal luminances, edge-stopped ----
    gx, gy = _depth_gradient(g)
    lt = luminance(colour)
    ws = np.zeros((h, w))
    s1 = np.zeros((h, w))
    win = []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            arg, fac = _edge_terms(P, g, shifted(g, dx, dy), dx, dy, gx, gy)
            wt = np.exp(-arg) * fac
            lq = shifted(lt, dx, dy)
            win.append((wt, lq))
            ws += wt
            s1 += wt * lq
    mean = s1 / np.where(ws > 0.0, ws, 1.0)
    s2 = sum(wt * (lq - mean) ** 2 for wt, lq in win)
    spatial = np.where(ws > 0.0, s2 / np.where(ws > 0.0, ws, 1.0), 0.0)
    cur = np.concatenate([colour, np.where(n_new < 4, spatial, var)[..., None]], -1)

    # ---- a-trous iterations ----
    history = None
    for k in range(P["iterations"]):
        s = 1 << k
        lc = luminance(cur)
        gv = sum(BIN3[i + 1] * BIN3[j + 1] * shifted(cur[..., 3], i * s, j * s)
                 for j in (-1, 0, 1) for i in (-1, 0, 1))
        sigma = P["phi_luma"] * np.sqrt(np.maximum(gv, 0.0) + P["variance_eps"])
        acc = np.zeros((h, w, 3))
        wsum = np.zeros((h, w))
        vsum = np.zeros((h, w))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                q = shifted(cur, dx * s, dy * s)
                arg, fac = _edge_terms(P, g, shifted(g, dx * s, dy * s), dx * s, dy * s, gx, gy)
                wt = B3[dx + 2] * B3[dy + 2] * np.exp(-np.abs(luminance(q) - lc) / sigma - arg) * fac
                acc += wt[..., None] * q[..., :3]
                wsum += wt
                vsum += wt * wt * q[..., 3]
        ok = wsum > 1e-6
        den = np.where(ok, wsum, 1.0)
        nxt = np.where(ok[..., None], np.concatenate([acc / den[..., None], (vsum / (den * den))[..., None]], -1), cur)
        if k == 0:
            history = np.concatenate([nxt[..., :3], n_new[..., None].astype(np.float64)], -1)
        cur = nxt
    denoised = np.concatenate([cur[..., :3], np.ones((h, w, 1))], -1)
    return history, mom, denoised
