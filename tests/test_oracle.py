"""Pins of the CPU oracle (CPU only).

The reference ships no test, golden vector or fixture for the shader, and Slang's CPU target is
not available here, so the oracle is pinned by what the mathematics fixes: closed forms of the
moment recursion, reductions to a separable B3-spline blur (Dammertz et al. 2010, cited at
svgf_denoise.slang:114), invariants (convexity, flip/transpose equivariance, constants), special
cases (NaN fallback, phi -> 0 identity), an independent brute-force restatement in Python loops,
and cross-checks between the oracle's independent code paths (float vs double, REFERENCE vs
CANONICAL with every canonical feature off, raw vs compensated moments).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O

LR, LG, LB = (float(np.float32(v)) for v in (0.2126, 0.7152, 0.0722))
K5 = [1 / 16, 4 / 16, 6 / 16, 4 / 16, 1 / 16]


def luma(img):
    return img[..., 0] * LR + img[..., 1] * LG + img[..., 2] * LB


def rand_img(rng, h, w, scale=10.0):
    img = (rng.random((h, w, 4)) * scale).astype(np.float32)
    img[..., 3] = 1.0
    return img


# ------------------------------------------------------------------------------------------
# binary16 conversion
# ------------------------------------------------------------------------------------------
def test_f16_rounding_matches_numpy():
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(1000.0),
        (rng.random(100000) * 1e-4).astype(np.float32),  # subnormal halves
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf, 2.0 ** -24, 2.0 ** -25,
                  2.0 ** -25 * 1.0000001, 6.1e-5, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11], np.float32),
    ])
    got = O.round_f16(vals)
    want = vals.astype(np.float16).astype(np.float32)
    assert np.array_equal(got, want)
    nan = O.round_f16(np.array([np.nan], np.float32))
    assert np.isnan(nan).all()


# ------------------------------------------------------------------------------------------
# REFERENCE mode: closed forms
# ------------------------------------------------------------------------------------------
def test_cold_first_frame_moments_closed_form():
    """From zero moments: mu1 = 0.2 l, mu2 = 0.2 l^2, var = 0.16 l^2, alpha = sqrt(var + 1e-4)."""
    rng = np.random.default_rng(2)
    L = rand_img(rng, 7, 9)
    mc, den, wa, wb = O.r_frame_f64(L.astype(np.float64), None, O.reference_params(1), history_valid=False) + (None, None)
    l = luma(L.astype(np.float64))
    a = float(np.float32(0.2))
    np.testing.assert_allclose(mc[..., 0], a * l, rtol=1e-14)
    np.testing.assert_allclose(mc[..., 1], a * l * l, rtol=1e-14)
    _, _, wa32, _ = O.r_frame_f32(L, None, O.reference_params(1), history_valid=False)
    want = np.sqrt(a * (1 - a) * l * l + float(np.float32(1e-4)))
    np.testing.assert_allclose(wa32[..., 3], want, rtol=2e-6)


def test_constant_luminance_k_frames_closed_form():
    """k frames of constant L from cold: mu1 = L(1-0.8^k), mu2 = L^2(1-0.8^k),
    var = L^2 (1-0.8^k) 0.8^k."""
    L = np.full((3, 4, 4), 3.0)
    l = 3.0 * (LR + LG + LB)
    a = float(np.float32(0.2))
    m = None
    for k in range(1, 9):
        m, _ = O.r_frame_f64(L, m, O.reference_params(1), history_valid=k > 1)
        q = (1 - a) ** k
        np.testing.assert_allclose(m[..., 0], l * (1 - q), rtol=1e-13)
        np.testing.assert_allclose(m[..., 1], l * l * (1 - q), rtol=1e-13)
        np.testing.assert_allclose(m[..., 1] - m[..., 0] ** 2, l * l * (1 - q) * q, rtol=1e-9)


def test_constant_image_is_fixed_point():
    L = np.zeros((6, 8, 4))
    L[..., 0], L[..., 1], L[..., 2] = 1.5, 2.5, 0.5
    _, den = O.r_frame_f64(L, None, O.reference_params(3), history_valid=False)
    np.testing.assert_allclose(den[..., :3], L[..., :3], rtol=1e-15)
    assert np.all(den[..., 3] == 1.0)
    _, den32, _, _ = O.r_frame_f32(L.astype(np.float32), None, O.reference_params(5), history_valid=False)
    np.testing.assert_allclose(den32[..., :3], L[..., :3], rtol=3e-7)


def _b3_blur_clamped(img, stride):
    """Separable [1,4,6,4,1]/16 at `stride` with clamp-to-edge, two 1-D passes (numpy)."""
    h, w = img.shape[:2]
    xs = np.arange(w)
    ys = np.arange(h)
    tmp = np.zeros_like(img)
    for k, d in zip(K5, range(-2, 3)):
        tmp += k * img[:, np.clip(xs + d * stride, 0, w - 1)]
    out = np.zeros_like(img)
    for k, d in zip(K5, range(-2, 3)):
        out += k * tmp[np.clip(ys + d * stride, 0, h - 1)]
    return out


def test_equal_luma_reduces_to_b3_spline_blur():
    """All pixels share one luminance (any chroma): every edge weight is exp(0) = 1, so the
    three iterations are the plain separable B3-spline a-trous blur at strides 1, 2, 4."""
    rng = np.random.default_rng(3)
    h, w = 13, 17
    r = rng.random((h, w)) * 4
    b = rng.random((h, w)) * 4
    g = (5.0 - LR * r - LB * b) / LG
    L = np.stack([r, g, b, np.ones_like(r)], -1)
    assert np.ptp(luma(L)) < 1e-13
    _, den = O.r_frame_f64(L, None, O.reference_params(3), history_valid=False)
    want = L[..., :3]
    for s in (1, 2, 4):
        want = _b3_blur_clamped(want, s)
    np.testing.assert_allclose(den[..., :3], want, rtol=1e-9, atol=1e-12)


def test_reference_luma_edge_stop_on_a_two_level_step():
    """A grey step (luminance la left of column 6, lb from it) through the cold first frame and
    one a-trous pass: every tap on the centre's side weighs K[dx], every tap across weighs
    K[dx] r with r = exp(-|lb - la| / max(2 sigma_c, eps)) and sigma_c = sqrt(0.16 l_c^2 + eps)
    the centre's (cold moments: var = alpha (1 - alpha) l^2). So the four pixels next to the
    step are (S_same l_c + S_cross r l_o) / (S_same + S_cross r) with S = 15/16, 1/16 or 11/16,
    5/16, and all others keep their level. Fixes the sign, the factor 2 and the epsilon of the
    edge stop, the centre's (not the tap's) sigma, and the normalisation."""
    w, h, edge = 12, 3, 6
    a, b = 1.0, 4.0
    L = np.zeros((h, w, 4))
    L[:, :edge, :3] = a
    L[:, edge:, :3] = b
    L[..., 3] = 1.0
    p = O.reference_params(1)
    _, den = O.r_frame_f64(L, None, p, history_valid=False)
    la, lb = luma(L[0, 0]), luma(L[0, -1])
    alpha, eps = float(np.float32(0.2)), float(np.float32(1e-4))

    def r(lc):
        sig = math.sqrt(alpha * (1 - alpha) * lc * lc + eps)
        return math.exp(-abs(lb - la) / max(2.0 * sig, eps))

    want = np.where(np.arange(w) < edge, a, b).astype(np.float64)
    for x, same, cross, lc, mine, other in ((4, 15 / 16, 1 / 16, la, a, b), (5, 11 / 16, 5 / 16, la, a, b),
                                            (6, 11 / 16, 5 / 16, lb, b, a), (7, 15 / 16, 1 / 16, lb, b, a)):
        want[x] = (same * mine + cross * r(lc) * other) / (same + cross * r(lc))
    for y in range(h):
        for c in range(3):
            np.testing.assert_allclose(den[y, :, c], want, rtol=1e-12)


def test_canonical_depth_and_normal_stops_are_separable_filters():
    """Constant luminance (any chroma), a depth ramp z = 4 + x / 128 and normals turning by
    theta per column, every row alike: then |z_q - z_p| = |dx| s / 128, the central-difference
    gradient is (1/128, 0) and n_p.n_q = cos(theta dx s), so a-trous iteration k is the fixed
    separable filter K[dy] x K[dx] exp(-|dx| s / (128 (phi_z |dx| s / 128 + eps_z)))
    cos(theta |dx| s)^phi_n, s = 2^k: the whole pass is five normalised 1-D convolutions along
    the rows (numpy). Compared away from the image edges. Fixes the depth term's gradient,
    stride, phi_z and epsilon, and the normal term's exponent."""
    w, h = 256, 4
    x = np.arange(w)
    theta = 0.003
    r = 1.0 + 0.5 * np.sin(0.3 * x)
    bb = 1.0 + 0.5 * np.cos(0.17 * x)
    g = (1.5 - LR * r - LB * bb) / LG
    rad = np.zeros((h, w, 4), np.float32)
    rad[..., 0], rad[..., 1], rad[..., 2], rad[..., 3] = r, g, bb, 1.0
    G = np.zeros((h, w, 4), np.float32)
    G[..., 0], G[..., 2], G[..., 3] = np.sin(theta * x), np.cos(theta * x), 4.0 + x / 128.0
    p = O.canonical_params(5)
    out = O.c_frame_f64(rad, G, None, np.zeros((h, w, 2), np.float32), None, None, p, history_valid=False)
    phi_z, eps_z, phi_n = float(p.phi_depth), float(p.depth_eps), float(p.phi_normal)
    c = rad[0, :, :3].astype(np.float64)
    for k in range(5):
        s = 1 << k
        wts = []
        for dx in range(-2, 3):
            t = abs(dx) * s / 128.0
            wts.append(K5[dx + 2] * math.exp(-t / (phi_z * t + eps_z)) * max(0.0, math.cos(theta * abs(dx) * s)) ** phi_n)
        c = sum(wt * c[np.clip(x + dx * s, 0, w - 1)] for wt, dx in zip(wts, range(-2, 3))) / sum(wts)
    inner = slice(64, 192)
    for y in range(h):
        np.testing.assert_allclose(out["denoised"][y, inner, :3], c[inner], rtol=1e-4)
    # the depth and normal terms matter here: a plain B3 blur differs by far more than the bound
    plain = rad[0, :, :3].astype(np.float64)
    for k in range(5):
        plain = _b3_blur_clamped(plain[None], 1 << k)[0]
    assert np.abs(plain[inner] - c[inner]).max() > 1e-2


def test_phi_zero_is_identity_where_luma_differs():
    """sigma -> 0 kills every tap whose luminance differs from the centre; interior pixels
    with distinct luminances come back unchanged."""
    rng = np.random.default_rng(4)
    L = rand_img(rng, 20, 20).astype(np.float64)
    p = O.reference_params(3)
    p.phi_luma = 0.0
    p.variance_eps = 1e-30
    _, den = O.r_frame_f64(L, None, p, history_valid=False)
    inner = np.s_[9:11, 9:11]  # stride-4 taps of these reach at most 8 px: no clamped duplicates
    np.testing.assert_array_equal(den[inner][..., :3], L[inner][..., :3])


# ------------------------------------------------------------------------------------------
# REFERENCE mode: invariants
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("op", ["flipx", "flipy", "transpose"])
def test_equivariance(op):
    rng = np.random.default_rng(5)
    L = rand_img(rng, 11, 11).astype(np.float64)
    mp = rng.random((11, 11, 2)) * 4
    f = {"flipx": lambda a: a[:, ::-1], "flipy": lambda a: a[::-1], "transpose": lambda a: a.transpose(1, 0, 2)}[op]
    m1, d1 = O.r_frame_f64(L, mp, O.reference_params(3))
    m2, d2 = O.r_frame_f64(np.ascontiguousarray(f(L)), np.ascontiguousarray(f(mp)), O.reference_params(3))
    np.testing.assert_allclose(f(d1), d2, rtol=1e-12)
    np.testing.assert_allclose(f(m1), m2, rtol=1e-14)


def test_convex_and_nonnegative():
    rng = np.random.default_rng(6)
    L = rand_img(rng, 16, 16, 100.0)
    _, den, _, _ = O.r_frame_f32(L, None, O.reference_params(3), history_valid=False)
    for c in range(3):
        assert den[..., c].min() >= L[..., c].min() * (1 - 1e-6)
        assert den[..., c].max() <= L[..., c].max() * (1 + 1e-6)
    assert np.all(den[..., 3] == 1.0)


def test_nan_input_falls_back_to_centre():
    """A NaN tap makes sum(w) NaN; `weight_sum > 1e-6` is false, so the centre colour is kept
    (svgf_denoise.slang:256). Pixels farther than 2 px from the NaN are unaffected at pass 1."""
    rng = np.random.default_rng(7)
    L = rand_img(rng, 9, 9)
    Ln = L.copy()
    Ln[4, 4, 1] = np.nan
    p = O.reference_params(1)
    _, den, wa, _ = O.r_frame_f32(Ln, None, p, history_valid=False)
    _, den_ok, _, _ = O.r_frame_f32(L, None, p, history_valid=False)
    assert np.array_equal(den[2, 3, :3], wa[2, 3, :3])  # within 2 px of the NaN: centre kept
    assert np.array_equal(den[0, 0], den_ok[0, 0])


# ------------------------------------------------------------------------------------------
# REFERENCE mode: independent brute force and cross-checks
# ------------------------------------------------------------------------------------------
def brute_force_reference(L, mp, iterations, phi=2.0, alpha=0.2, eps=1e-4):
    """Pure-Python loops over svgf_denoise.slang:152-275 (written apart from the C oracle)."""
    h, w = len(L), len(L[0])
    alpha, eps, phi = float(np.float32(alpha)), float(np.float32(eps)), float(phi)
    lum = lambda c: c[0] * LR + c[1] * LG + c[2] * LB  # noqa: E731
    A = [[None] * w for _ in range(h)]
    M = [[None] * w for _ in range(h)]
    for y in range(h):
        for x in range(w):
            c = L[y][x]
            l = lum(c)
            m1, m2 = mp[y][x] if mp is not None else (0.0, 0.0)
            n1 = m1 + alpha * (l - m1)
            n2 = m2 + alpha * (l * l - m2)
            M[y][x] = (n1, n2)
            A[y][x] = (c[0], c[1], c[2], math.sqrt(max(n2 - n1 * n1, 0.0) + eps))
    B = [[None] * w for _ in range(h)]
    out = None
    for p in range(1, iterations + 1):
        st = 1 << (p - 1)
        src, dst = (A, B) if p % 2 == 1 else (B, A)
        res = [[None] * w for _ in range(h)]
        for y in range(h):
            for x in range(w):
                c = src[y][x]
                lc = lum(c)
                sg = max(phi * c[3], eps)
                acc = [0.0, 0.0, 0.0]
                ws = 0.0
                for dy in range(-2, 3):
                    for dx in range(-2, 3):
                        q = src[min(max(y + dy * st, 0), h - 1)][min(max(x + dx * st, 0), w - 1)]
                        wt = K5[dx + 2] * K5[dy + 2] * math.exp(-abs(lum(q) - lc) / sg)
                        for i in range(3):
                            acc[i] += q[i] * wt
                        ws += wt
                o = [a / ws for a in acc] if ws > 1e-6 else list(c[:3])
                res[y][x] = (o[0], o[1], o[2], 1.0 if p == iterations else c[3])
        for y in range(h):
            for x in range(w):
                dst[y][x] = res[y][x]
        out = res
    return M, out


@pytest.mark.parametrize("shape,iters", [((6, 8), 3), ((7, 13), 3), ((1, 1), 3), ((9, 2), 5), ((2, 9), 4)])
def test_brute_force_matches_oracle(shape, iters):
    rng = np.random.default_rng(8)
    h, w = shape
    L = rand_img(rng, h, w).astype(np.float64)
    mp = rng.random((h, w, 2)) * 5
    M, D = brute_force_reference(L.tolist(), mp.tolist(), iters)
    mc, den = O.r_frame_f64(L, mp, O.reference_params(iters))
    np.testing.assert_allclose(mc, np.array(M), rtol=1e-13)
    np.testing.assert_allclose(den, np.array(D), rtol=1e-12)


def test_float_oracle_close_to_double():
    rng = np.random.default_rng(9)
    L = rand_img(rng, 40, 50)
    mp = (rng.random((40, 50, 2)) * 5).astype(np.float32)
    m32, d32, _, _ = O.r_frame_f32(L, mp, O.reference_params(3))
    m64, d64 = O.r_frame_f64(L.astype(np.float64), mp.astype(np.float64), O.reference_params(3))
    # The shader's variance max(mu2 - mu1^2, 0) cancels catastrophically where the moments are
    # nearly consistent; sigma = sqrt(var + 1e-4) then carries the float rounding of mu2 into
    # the edge stop. Float and double agree on average, not to the last digits everywhere.
    err = np.abs(d32 - d64) / np.abs(d64).max()
    assert err.max() < 5e-5 and err.mean() < 5e-7


def test_threads_do_not_change_result():
    rng = np.random.default_rng(10)
    L = rand_img(rng, 37, 29)
    a = O.r_frame_f32(L, None, O.reference_params(3), history_valid=False, threads=1)
    b = O.r_frame_f32(L, None, O.reference_params(3), history_valid=False, threads=5)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_half_store_variant_quantises_every_store():
    rng = np.random.default_rng(11)
    L = O.round_f16(rand_img(rng, 10, 12))
    m, d, wa, wb = O.r_frame_f32(L, None, O.reference_params(3), store=O.STORE_F16, history_valid=False)
    for a in (m, d, wa, wb):
        assert np.array_equal(a, a.astype(np.float16).astype(np.float32))


def test_half_store_moments_overflow_to_inf():
    """Luminance above 255.9 overflows the rg16f second moment (the hazard of the reference's
    own format): mu2 = 0.2 l^2 > 65504 stores +inf."""
    L = np.full((2, 2, 4), 600.0, np.float32)
    m, _, _, _ = O.r_frame_f32(L, None, O.reference_params(1), store=O.STORE_F16, history_valid=False)
    assert np.isinf(m[..., 1]).all() and np.isfinite(m[..., 0]).all()


# ------------------------------------------------------------------------------------------
# CANONICAL oracle
# ------------------------------------------------------------------------------------------
def _as_reference_params(iters=3):
    return O.OracleParams(iters, 2.0, 0.0, 0.0, 1.0, 0.2, 1e-4, 1e-3, 0.0, 0.0, 1, O.FLAG_AS_REFERENCE)


def test_canonical_with_features_off_equals_reference():
    """Same-pixel moments, alpha carries sigma, luminance-only edge stop, no variance filter:
    the CANONICAL code path reproduces the separately written REFERENCE restatement."""
    rng = np.random.default_rng(12)
    L = rand_img(rng, 15, 21)
    mp = rng.random((15, 21, 2)) * 5
    for hv in (False, True):
        mc, den = O.r_frame_f64(L.astype(np.float64), mp, O.reference_params(3), history_valid=hv)
        out = O.c_frame_f64(L, None, None, None, None, mp, _as_reference_params(3), history_valid=hv)
        np.testing.assert_allclose(out["denoised"], den, rtol=1e-13)
        np.testing.assert_allclose(out["moments"], mc, rtol=1e-14)


def _flat_g(h, w, z=5.0):
    g = np.zeros((h, w, 4), np.float32)
    g[..., 2] = 1.0
    g[..., 3] = z
    return g


def test_canonical_cold_equals_full_disocclusion():
    """History whose depth fails the test everywhere gives the cold-start result."""
    rng = np.random.default_rng(13)
    h, w = 12, 16
    L = rand_img(rng, h, w)
    g = _flat_g(h, w)
    gp = _flat_g(h, w, 50.0)
    mv = np.full((h, w, 2), 0.3, np.float32)
    hist = rng.random((h, w, 4)) * 3
    hist[..., 3] = 7
    mom = rng.random((h, w, 2)) * 3
    cold = O.c_frame_f64(L, g, None, mv, None, None, O.canonical_params(), history_valid=False, debug=True)
    dis = O.c_frame_f64(L, g, gp, mv, hist, mom, O.canonical_params(), history_valid=True, debug=True)
    for k in ("denoised", "history", "moments", "temporal"):
        np.testing.assert_array_equal(cold[k], dis[k])
    assert np.all(cold["temporal"][..., 3] == 1)


@pytest.mark.parametrize("shift", [0, 1])
def test_canonical_integer_pan_reprojects_exactly(shift):
    """Motion (+shift, 0) with a consistent G-buffer: the temporal colour is
    H[y, x+shift] + a (R - H[y, x+shift]) with a = max(0.2, 1/n), n = n_hist + 1. Zero motion is
    same-pixel history (the reference's own temporal rule, which has no reprojection)."""
    rng = np.random.default_rng(14)
    h, w = 6, 10
    L = rand_img(rng, h, w)
    g = _flat_g(h, w)
    mv = np.zeros((h, w, 2), np.float32)
    mv[..., 0] = shift
    hist = rng.random((h, w, 4)) * 3
    hist[..., 3] = np.array([1, 2, 3, 9, 40, 1, 2, 3, 4, 5])[None, :]
    mom = rng.random((h, w, 2)) * 3
    p = O.canonical_params()
    out = O.c_frame_f64(L, g, g, mv, hist, mom, p, history_valid=True, debug=True)
    t = out["temporal"]
    for y in range(h):
        for x in range(w - shift):
            n = min(int(hist[y, x + shift, 3]) + 1, 32)
            a = max(float(np.float32(0.2)), 1.0 / n)
            H = hist[y, x + shift, :3]
            np.testing.assert_allclose(t[y, x, :3], H + a * (L[y, x, :3] - H), rtol=1e-14)
            assert t[y, x, 3] == n
        if shift:
            assert t[y, w - 1, 3] == 1  # reprojects outside the image: disoccluded


def test_canonical_raw_and_compensated_moments_agree():
    """The West/Finch recursion on (mean, centred second moment), with the exact bilinear
    interpolation of the centred state, equals the raw E[x^2] - E[x]^2 form in exact arithmetic;
    in double they agree to rounding."""
    rng = np.random.default_rng(15)
    h, w = 14, 18
    g = _flat_g(h, w)
    g[..., 3] = 5 + rng.random((h, w)).astype(np.float32) * 0.01
    seq = [(rand_img(rng, h, w), (rng.random((h, w, 2)) * 2 - 1).astype(np.float32)) for _ in range(5)]
    outs = {}
    for comp in (False, True):
        p = O.canonical_params(5, compensated=comp)
        hist = mom = None
        for f, (L, mv) in enumerate(seq):
            o = O.c_frame_f64(L, g, g, mv, hist, mom, p, history_valid=f > 0)
            hist, mom = o["history"], o["moments"]
        outs[comp] = o
    np.testing.assert_allclose(outs[True]["denoised"], outs[False]["denoised"], rtol=1e-9)
    var_raw = outs[False]["moments"][..., 1] - outs[False]["moments"][..., 0] ** 2
    np.testing.assert_allclose(outs[True]["moments"][..., 1] ** 2, var_raw, rtol=1e-6, atol=1e-9)


def test_canonical_flat_edges_reduce_to_b3_blur_of_variance():
    """phi_n = phi_z = 0 and equal luminance: each iteration is the clamped B3 blur of the
    colour and the K^2-weighted blur of the variance."""
    rng = np.random.default_rng(16)
    h, w = 11, 13
    r = rng.random((h, w)) * 4
    b = rng.random((h, w)) * 4
    gg = (5.0 - LR * r - LB * b) / LG
    L = np.stack([r, gg, b, np.ones_like(r)], -1)
    p = O.canonical_params(2)
    p.phi_normal = 0.0
    p.phi_depth = 0.0
    g = _flat_g(h, w)
    out = O.c_frame_f64(L.astype(np.float32), g, None, np.zeros((h, w, 2), np.float32), None, None, p,
                        history_valid=False, debug=True)
    W0 = out["variance"]
    col = W0[..., :3]
    for s in (1, 2):
        col = _b3_blur_clamped(col, s)
    np.testing.assert_allclose(out["denoised"][..., :3], col, rtol=1e-6)


def brute_force_canonical_atrous(W0, G, p, iterations):
    """Pure-Python loops over the CANONICAL a-trous iterations (written apart from the C
    oracle): lattice 3x3 binomial variance prefilter, per-pixel depth gradient, depth /
    normal / luminance edge stops, variance filtered by squared weights."""
    h, w = len(W0), len(W0[0])
    lum = lambda c: c[0] * LR + c[1] * LG + c[2] * LB  # noqa: E731
    cl = lambda v, hi: min(max(v, 0), hi)  # noqa: E731
    k3 = [0.25, 0.5, 0.25]
    src = [[tuple(W0[y][x]) for x in range(w)] for y in range(h)]
    hist = None
    for k in range(iterations):
        s = 1 << k
        dst = [[None] * w for _ in range(h)]
        for y in range(h):
            for x in range(w):
                c = src[y][x]
                gv = sum(k3[i + 1] * k3[j + 1] * src[cl(y + j * s, h - 1)][cl(x + i * s, w - 1)][3]
                         for j in (-1, 0, 1) for i in (-1, 0, 1))
                sig = float(p.phi_luma) * math.sqrt(max(gv, 0.0) + float(p.variance_eps))
                gz = lambda xx, yy: float(G[cl(yy, h - 1)][cl(xx, w - 1)][3])  # noqa: E731
                gx = 0.5 * (gz(x + 1, y) - gz(x - 1, y))
                gy = 0.5 * (gz(x, y + 1) - gz(x, y - 1))
                gp = G[y][x]
                acc = [0.0, 0.0, 0.0]
                sw = sv = 0.0
                for dy in range(-2, 3):
                    for dx in range(-2, 3):
                        qx, qy = cl(x + dx * s, w - 1), cl(y + dy * s, h - 1)
                        q = src[qy][qx]
                        gq = G[qy][qx]
                        a = abs(lum(q) - lum(c)) / sig
                        a += abs(float(gq[3]) - float(gp[3])) / (
                            float(p.phi_depth) * abs(gx * dx * s + gy * dy * s) + float(p.depth_eps))
                        d = sum(float(gp[t]) * float(gq[t]) for t in range(3))
                        wn = d ** float(p.phi_normal) if d > 0 else 0.0
                        wt = K5[dx + 2] * K5[dy + 2] * math.exp(-a) * wn
                        for t in range(3):
                            acc[t] += wt * q[t]
                        sw += wt
                        sv += wt * wt * q[3]
                dst[y][x] = (acc[0] / sw, acc[1] / sw, acc[2] / sw, sv / (sw * sw))
        if k == 0:
            hist = dst
        src = dst
    return np.array(src), np.array(hist)


def test_canonical_atrous_brute_force():
    rng = np.random.default_rng(19)
    h, w = 9, 11
    L = rand_img(rng, h, w)
    g = np.zeros((h, w, 4), np.float32)
    n = rng.standard_normal((h, w, 3)) * 0.15 + np.array([0, 0, 1])
    g[..., :3] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    g[..., 3] = 4 + rng.random((h, w)).astype(np.float32)
    p = O.canonical_params(3)
    p.phi_normal = 8.0  # weaker normal stop so every term matters
    out = O.c_frame_f64(L, g, None, np.zeros((h, w, 2), np.float32), None, None, p, history_valid=False, debug=True)
    want, hist = brute_force_canonical_atrous(out["variance"].tolist(), g.tolist(), p, 3)
    np.testing.assert_allclose(out["denoised"][..., :3], want[..., :3], rtol=1e-11)
    np.testing.assert_allclose(out["history"][..., :3], hist[..., :3], rtol=1e-11)


def test_canonical_threads_do_not_change_result():
    rng = np.random.default_rng(18)
    h, w = 23, 31
    L = rand_img(rng, h, w)
    g = _flat_g(h, w)
    g[..., 3] = 3 + rng.random((h, w)).astype(np.float32)
    mv = (rng.random((h, w, 2)) * 2 - 1).astype(np.float32)
    hist = rng.random((h, w, 4)) * 3
    hist[..., 3] = 3
    mom = rng.random((h, w, 2))
    a = O.c_frame_f64(L, g, g, mv, hist, mom, O.canonical_params(), debug=True, threads=1)
    b = O.c_frame_f64(L, g, g, mv, hist, mom, O.canonical_params(), debug=True, threads=6)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_canonical_spatial_variance_is_window_variance():
    """Young history (n < 4) with depth / normal stops off: the variance is the plain
    3x3 (clamped) variance of the temporal luminance."""
    rng = np.random.default_rng(17)
    h, w = 7, 9
    L = rand_img(rng, h, w)
    p = O.canonical_params(1)
    p.phi_normal = 0.0
    p.phi_depth = 0.0
    out = O.c_frame_f64(L, _flat_g(h, w), None, np.zeros((h, w, 2), np.float32), None, None, p,
                        history_valid=False, debug=True)
    lum = luma(L.astype(np.float64))
    for y in range(h):
        for x in range(w):
            win = [lum[min(max(y + dy, 0), h - 1), min(max(x + dx, 0), w - 1)] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
            np.testing.assert_allclose(out["variance"][y, x, 3], np.var(win), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("noise", [0, 1], ids=["exponential", "lognormal"])
def test_canonical_deviations_keep_the_denoising_quality(noise):
    """CANONICAL deviates from Schied et al. 2017 in two places: its 3x3 variance prefilter reads
    the pixel's lattice neighbours at +-s (Schied: +-1), and its luminance scale is
    phi_l sqrt(g(var) + eps) (Schied: phi_l sqrt(g(var)) + eps). Against the noise-free radiance
    of the same sequence (12 frames, pan, 10% disocclusion), the relMSE of the output with this
    build's forms stays within 3% of Schied's forms' (both 8-15x below the input's): the
    deviations move individual pixels, not the quality."""
    import svgf_amd as S
    from svgf_testutil import oracle_canonical_sequence, seq_inputs

    S.lib()
    w, h, frames = 320, 180, 12
    ins = seq_inputs(w, h, frames, noise=noise, pan_px=1.37, disocclusion=0.1)
    gt = [S.synth_frame(w, h, seed=f, frame=f, noise=2, pan_px=1.37, disocclusion=0.1)["radiance"][..., :3]
          .astype(np.float64) for f in range(frames)]

    def relmse(o, g):
        return float(np.mean((o - g) ** 2 / (g ** 2 + 1e-2)))

    q = {}
    for name, flags in (("build", 0), ("schied", O.FLAG_PREFILTER_UNIT | O.FLAG_EPS_OUTSIDE)):
        p = O.canonical_params(5, compensated=True)
        p.flags |= flags
        outs = oracle_canonical_sequence(ins, p, threads=8)
        q[name] = np.mean([relmse(o["denoised"][..., :3], g) for o, g in zip(outs[6:], gt[6:])])
    noisy = np.mean([relmse(d["radiance"][..., :3].astype(np.float64), g) for d, g in zip(ins[6:], gt[6:])])
    print(f"noise {noise}: relMSE input {noisy:.4f}, this build {q['build']:.5f}, Schied's forms {q['schied']:.5f}")
    assert abs(q["build"] / q["schied"] - 1.0) <= 0.03
    assert q["build"] < noisy / 8.0


@pytest.mark.parametrize("compensated", [False, True])
def test_canonical_crop_interior_equals_full_frame(compensated):
    """Locality of the CANONICAL pass: over 3 frames with fractional pan and disocclusion, the
    oracle on a 448x448 crop reproduces the full-frame oracle exactly inside the crop's
    interior (MARGIN from every edge that is not an image edge). This pins the margin that the
    full-size GPU window test relies on; a margin below the receptive field fails here."""
    import svgf_amd as S
    from svgf_testutil import MARGIN, crop_interior, crop_windows, oracle_canonical_sequence, seq_inputs

    S.lib()
    w, h, c = 640, 520, 448
    inputs = seq_inputs(w, h, 3, noise=1, pan_px=1.37, disocclusion=0.1)
    op = O.canonical_params(5, compensated=compensated)
    full = oracle_canonical_sequence(inputs, op, threads=8)[-1]
    for y0, x0 in crop_windows(w, h, c):
        crop = [{k: np.ascontiguousarray(v[y0:y0 + c, x0:x0 + c]) for k, v in d.items()} for d in inputs]
        ref = oracle_canonical_sequence(crop, op, threads=8)[-1]
        ys, xs = crop_interior(y0, x0, c, w, h)
        for k in ("denoised", "history", "moments"):
            assert np.array_equal(full[k][y0:y0 + c, x0:x0 + c][ys, xs], ref[k][ys, xs]), (k, y0, x0)
        # the margin is needed: just outside the interior the crop edge changes the result
        if x0 > 0:
            a = full["denoised"][y0:y0 + c, x0:x0 + c][ys, :MARGIN // 4]
            assert not np.array_equal(a, ref["denoised"][ys, :MARGIN // 4])


# ------------------------------------------------------------------------------------------
# agent-frame wire encoder (oracle/wire_oracle.c)
# ------------------------------------------------------------------------------------------
def _srgb64(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.0031308, x * 12.92, 1.055 * np.power(np.maximum(x, 0.0), 1.0 / 2.4) - 0.055)


def test_wire_thresholds_are_the_smallest_floats_reaching_each_code():
    t = O.wire_thresholds()
    assert np.all(np.diff(t) > 0)
    k = np.arange(1, 256, dtype=np.float64)
    target = (k - 0.5) / 255.0
    below = np.nextafter(t, np.float32(0.0))
    assert np.all(_srgb64(t) >= target)
    assert np.all(_srgb64(below) < target)


def test_wire_codes_equal_rounded_srgb():
    """code(x) = floor(255 srgb(x) + 1/2) in double, on random values, every threshold and its
    float neighbours, 0, 1 and values outside [0, 1]."""
    rng = np.random.default_rng(7)
    t = O.wire_thresholds()
    x = np.concatenate([rng.random(200_000, dtype=np.float32), t, np.nextafter(t, np.float32(0)),
                        np.nextafter(t, np.float32(2)), np.float32([0.0, 1.0, 1e-9, 0.0031308, 0.5])])
    want = np.floor(255.0 * _srgb64(np.clip(x, 0, 1)) + 0.5).astype(np.uint8)
    assert np.array_equal(O.wire_codes(x), want)
    assert list(O.wire_codes(np.float32([-1.0, 2.0, np.inf]))) == [0, 255, 255]


def test_wire_tonemap_grey_axis_is_the_fitted_curve():
    """The input matrix has unit row sums (postprocess.slang:49-53), so a grey input is the same
    grey in AP1; output channel r is then the RRT+ODT fit of the grey level times row r's sum of
    the output matrix (1, 1 and 0.99999 for the published coefficients, :56-60). A transposed
    matrix (column sums 0.70 / 1.40 / 0.92 in, 1.50 / 0.50 / 1.00 out) or a wrong fit
    coefficient breaks this."""
    def fit(v):
        return (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)

    out_row_sums = np.array([1.60475 - 0.53108 - 0.07367, -0.10208 + 1.10813 - 0.00605, -0.00327 - 0.07276 + 1.07602])
    for g in [0.0, 1e-3, 0.01, 0.18, 0.5, 1.0, 2.0, 8.0, 64.0, 1e4]:
        out = O.wire_tonemap([g, g, g])
        want = np.clip(fit(g) * out_row_sums, 0.0, 1.0)
        assert np.allclose(out, want, rtol=2e-6, atol=2e-7), (g, out, want)
    # a saturated primary stays within [0, 1]; exposure scales the input
    assert np.all((O.wire_tonemap([50.0, 0.0, 0.0]) >= 0) & (O.wire_tonemap([50.0, 0.0, 0.0]) <= 1))
    assert np.array_equal(O.wire_tonemap([0.1, 0.2, 0.3], exposure=2.0), O.wire_tonemap([0.2, 0.4, 0.6]))


def test_wire_tonemap_is_monotonic_along_grey():
    g = np.float32(np.geomspace(1e-4, 1e3, 400))
    y = np.array([O.wire_tonemap([v, v, v])[0] for v in g])
    assert np.all(np.diff(y) >= 0)


def test_wire_depth_normalisation():
    near, far = O.WIRE_NEAR, O.WIRE_FAR
    L = O.lib()
    assert L.oracle_wire_depth(np.float32(near), near, far) == 0.0
    assert L.oracle_wire_depth(np.float32(far), near, far) == 1.0
    assert L.oracle_wire_depth(np.float32(500.0), near, far) == 1.0
    assert L.oracle_wire_depth(np.float32(0.05), near, far) == 0.0
    for z in (0.0, -3.0, float("nan"), float("inf")):
        assert L.oracle_wire_depth(z, near, far) == 1.0  # no geometry reads as far
    assert abs(L.oracle_wire_depth(np.float32(50.05), near, far) - 0.5) < 1e-6


def test_wire_encode_layout():
    """rgb[i] of pixel (x, y) sits at byte 3 (y w + x): rows top to bottom, no padding."""
    rng = np.random.default_rng(1)
    h, w = 5, 7
    col = rng.random((h, w, 4), dtype=np.float32) * 3
    g = np.zeros((h, w, 4), np.float32)
    g[..., 3] = 1.0 + rng.random((h, w), dtype=np.float32) * 50
    rgb, depth = O.wire_encode(col, g)
    for y, x in [(0, 0), (2, 3), (4, 6)]:
        assert np.array_equal(rgb[y, x], O.wire_codes(O.wire_tonemap(col[y, x, :3])))
        assert depth[y, x] == O.lib().oracle_wire_depth(g[y, x, 3], O.WIRE_NEAR, O.WIRE_FAR)
    assert rgb.tobytes()[3 * (2 * w + 3):3 * (2 * w + 3) + 3] == rgb[2, 3].tobytes()
