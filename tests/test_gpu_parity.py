"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same seeded inputs.

REFERENCE mode vs the float restatement of src/svgf_denoise.slang: <= 1e-5 relative error
(FP32 storage). FP16 storage vs the half-store restatement: within one half ulp. CANONICAL mode
vs the FP64 oracle: tolerances stated per test (FP32 arithmetic against binary64).
"""
from __future__ import annotations

import numpy as np
import pytest

from svgf_testutil import gpu_canonical_sequence, moment_errors, oracle_canonical_sequence, rel_err, seq_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import svgf_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    svgf_amd.lib()
    return svgf_amd


def _ref_gpu(S, ctx, lighting, moments_prev, dtype, history_valid):
    dev = torch.device("cuda")
    h, w = lighting.shape[:2]
    rad = torch.from_numpy(lighting).to(dev).to(dtype)
    den = torch.empty((h, w, 4), dtype=dtype, device=dev)
    mom = torch.empty((h, w, 2), dtype=dtype, device=dev)
    mp = torch.from_numpy(moments_prev).to(dev).to(dtype) if moments_prev is not None else None
    ctx.denoise(radiance=rad, denoised=den, moments_curr=mom, moments_prev=mp, history_valid=history_valid)
    torch.cuda.synchronize()
    return den.float().cpu().numpy(), mom.float().cpu().numpy()


@pytest.mark.parametrize("iterations", [1, 3, 5, 10])
def test_reference_fp32_640x360_warm_frame(S, oracle_lib, iterations):
    """640x360: history = a frame from seed 1, current frame from seed 0 (alpha 0.2)."""
    O = oracle_lib
    w, h = 640, 360
    prev = S.synth_frame(w, h, seed=1, frame=0, pan_px=0.5)["radiance"]
    cur = S.synth_frame(w, h, seed=0, frame=1, pan_px=0.5)["radiance"]
    p = S.default_params(S.MODE_REFERENCE)
    p.iterations = iterations
    op = O.reference_params(iterations)
    with S.Context(S.MODE_REFERENCE, S.STORAGE_FP32, params=p, max_width=w, max_height=h) as ctx:
        d0, m0 = _ref_gpu(S, ctx, prev, None, torch.float32, False)
        d1, m1 = _ref_gpu(S, ctx, cur, m0, torch.float32, True)
    om0, od0, _, _ = O.r_frame_f32(prev, None, op, history_valid=False)
    om1, od1, _, _ = O.r_frame_f32(cur, om0, op, history_valid=True)
    for a, b, name in ((d0, od0, "denoised[0]"), (m0, om0, "moments[0]"), (d1, od1, "denoised[1]"),
                       (m1, om1, "moments[1]")):
        mx, mean = rel_err(a, b)
        print(f"{name}: max rel {mx:.3e} mean {mean:.3e}")
        assert mx <= 1e-5, name
    assert np.all(d1[..., 3] == 1.0)


def test_reference_fp16_storage(S, oracle_lib):
    """The Vulkan rgba16f/rg16f form vs the restatement that rounds every store to half."""
    O = oracle_lib
    w, h = 320, 180
    prev = O.round_f16(S.synth_frame(w, h, seed=1, frame=0)["radiance"])
    cur = O.round_f16(S.synth_frame(w, h, seed=0, frame=1)["radiance"])
    op = O.reference_params(3)
    with S.Context(S.MODE_REFERENCE, S.STORAGE_FP16, max_width=w, max_height=h) as ctx:
        d0, m0 = _ref_gpu(S, ctx, prev, None, torch.float16, False)
        d1, m1 = _ref_gpu(S, ctx, cur, m0, torch.float16, True)
    om0, od0, _, _ = O.r_frame_f32(prev, None, op, store=O.STORE_F16, history_valid=False)
    om1, od1, _, _ = O.r_frame_f32(cur, om0, op, store=O.STORE_F16, history_valid=True)
    for a, b, name in ((d0, od0, "denoised[0]"), (m0, om0, "moments[0]"), (d1, od1, "denoised[1]"),
                       (m1, om1, "moments[1]")):
        mx, mean = rel_err(a, b)
        print(f"{name}: max rel {mx:.3e} mean {mean:.3e}")
        assert mx <= 2.0 ** -10 + 1e-6, name  # one half ulp of disagreement on a rounding tie
        assert mean <= 1e-5, name


def test_reference_fp16_moment_overflow_and_nan_taps(S, oracle_lib):
    """The reference's own FP16 hazards, against the restatement that rounds every store to half:
    luminance above 256 makes the rg16f second moment overflow to +inf (then NaN through the
    lerp a + t (b - a) on the next frame, and variance max(NaN - mu^2, 0) = 0); a NaN radiance
    tap makes the weight sum NaN, so `weight_sum > 1e-6` is false and the pixel keeps its
    centre colour (svgf_denoise.slang:184-189, 256; main.cpp:1367-1370). 3 frames: every output
    within one half ulp, non-finite values in the same places, moments bit for bit."""
    O = oracle_lib
    w, h = 320, 180
    rng = np.random.default_rng(9)
    seq = []
    for f in range(3):
        L = S.synth_frame(w, h, seed=40 + f, frame=f)["radiance"].copy()
        L[40:120, 60:200, :3] *= 200.0  # luminance up to ~2000: mu2 = 0.2 l^2 overflows binary16
        if f == 1:
            ys, xs = rng.integers(0, h, 25), rng.integers(0, w, 25)
            L[ys, xs, rng.integers(0, 3, 25)] = np.nan
        seq.append(O.round_f16(L))
    op = O.reference_params(3)
    dev = torch.device("cuda")
    mom = [torch.zeros((h, w, 2), dtype=torch.float16, device=dev) for _ in range(2)]
    om = None
    with S.Context(S.MODE_REFERENCE, S.STORAGE_FP16, max_width=w, max_height=h) as ctx:
        for f, L in enumerate(seq):
            den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
            ctx.denoise(radiance=torch.from_numpy(L).to(dev).half(), denoised=den, moments_curr=mom[f % 2],
                        moments_prev=mom[(f + 1) % 2] if f else None, history_valid=f > 0)
            torch.cuda.synchronize()
            om, od, _, _ = O.r_frame_f32(L, om, op, store=O.STORE_F16, history_valid=f > 0)
            g = den.float().cpu().numpy()
            gm = mom[f % 2].float().cpu().numpy()
            assert np.isinf(om[..., 1]).any(), f  # the hazard is exercised
            if f == 2:
                assert np.isnan(om[..., 1]).any()
            assert np.array_equal(gm, om, equal_nan=True), f
            fin = np.isfinite(od)
            assert np.array_equal(np.isfinite(g), fin) and np.array_equal(np.isnan(g), np.isnan(od)), f
            mx, _ = rel_err(g[fin], od[fin])
            print(f"frame {f}: {int((~fin).sum())} non-finite outputs, max rel err {mx:.3e}")
            assert mx <= 2.0 ** -10 + 1e-6, f


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("shape", [(1, 1), (2, 9), (9, 2), (65, 5), (67, 131)])
def test_reference_edge_sizes(S, oracle_lib, shape, storage):
    """Frames smaller than the taps' reach (every tap clamped), a cold and a warm frame: FP32
    within 1e-5 of the float restatement; FP16 storage within one half ulp of the half-store
    restatement, moments bit for bit."""
    O = oracle_lib
    w, h = shape
    fp16 = storage == "fp16"
    st, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    seq = [S.synth_frame(w, h, seed=3 + f, frame=f)["radiance"] for f in range(2)]
    if fp16:
        seq = [O.round_f16(L) for L in seq]
    om = None
    with S.Context(S.MODE_REFERENCE, st, max_width=w, max_height=h) as ctx:
        m = None
        for f, L in enumerate(seq):
            d, m = _ref_gpu(S, ctx, L, m, dt, f > 0)
            om, od, _, _ = O.r_frame_f32(L, om, O.reference_params(3), store=O.STORE_F16 if fp16 else O.STORE_F32,
                                         history_valid=f > 0)
            if fp16:
                assert rel_err(d, od)[0] <= 2.0 ** -10 + 1e-6, f
                assert np.array_equal(m, om), f
            else:
                assert rel_err(d, od)[0] <= 1e-5, f
                assert rel_err(m, om)[0] <= 1e-5, f


def test_canonical_fp32_sequence(S, oracle_lib):
    """6 frames, fractional pan, 10% forced disocclusion: FP32 GPU vs FP64 oracle."""
    O = oracle_lib
    w, h = 200, 120
    inputs = seq_inputs(w, h, 6, pan_px=1.37, disocclusion=0.1)
    op = O.canonical_params(5)
    ref = oracle_canonical_sequence(inputs, op)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP32, max_width=w, max_height=h) as ctx:
        got = gpu_canonical_sequence(ctx, inputs, torch.float32)
    for f, (g, r) in enumerate(zip(got, ref)):
        dm, dmean = rel_err(g["denoised"][..., :3], r["denoised"][..., :3])
        hm, _ = rel_err(g["history"], r["history"])
        print(f"frame {f}: denoised max {dm:.3e} mean {dmean:.3e}; history max {hm:.3e}")
        assert dm <= 2e-4 and dmean <= 1e-6, f
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), "history length differs"


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
@pytest.mark.parametrize("shape", [(1, 1), (2, 9), (9, 2), (17, 3), (65, 5), (33, 257), (67, 131)])
def test_canonical_edge_sizes(S, oracle_lib, shape, form):
    """Frames narrower or shorter than the a-trous strides (lattice residues with no pixel,
    single-point lattices, tiles cut by both edges), 4 frames each, vs the FP64 oracle: FP32
    within 2e-4, the FP16 wide form within 1e-3 max relative error."""
    O = oracle_lib
    w, h = shape
    fp16 = form == "fp16_wide"
    inputs = seq_inputs(w, h, 4, seed0=50, pan_px=0.9, disocclusion=0.1, half=fp16)
    ref = oracle_canonical_sequence(inputs, O.canonical_params(5, compensated=fp16))
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    bound = 1e-3 if fp16 else 2e-4
    with S.Context(S.MODE_CANONICAL, storage, max_width=w, max_height=h) as ctx:
        got = gpu_canonical_sequence(ctx, inputs, dt)
    for f, (g, r) in enumerate(zip(got, ref)):
        assert np.isfinite(g["denoised"]).all(), f
        assert rel_err(g["denoised"][..., :3], r["denoised"][..., :3])[0] <= bound, f
        assert rel_err(g["history"][..., :3], r["history"][..., :3])[0] <= bound, f
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        if fp16:
            me = moment_errors(g["moments"], r["moments"])
            assert me["mean"][0] <= bound and me["second"][0] <= bound, f


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
def test_canonical_incoherent_motion(S, oracle_lib, form):
    """The temporal stage reads a pixel's four history taps from the tile's staged footprint
    (the tile moved by its centre pixel's motion) when they lie in it, and gathers them
    otherwise. Motion fields that mix both within tiles: per-pixel jitter of up to +-2.5 px, bands
    of a far coherent shift (-7.3 px), and pixels with the no-motion sentinel (-1e6) or motion
    off the frame. 4 frames against the FP64 oracle, history lengths exact."""
    O = oracle_lib
    w, h = 150, 90
    fp16 = form != "fp32"
    inputs = seq_inputs(w, h, 4, noise=1 if fp16 else 0, pan_px=1.37, disocclusion=0.1, half=fp16)
    rng = np.random.default_rng(17)
    for f, d in enumerate(inputs):
        if f == 0:
            continue
        m = d["motion"].copy()
        m += rng.uniform(-2.5, 2.5, m.shape).astype(np.float32)
        m[20:40, :, 0] = -7.3
        m[:, 100:120, 1] = 5.6
        m[rng.random((h, w)) < 0.02] = -1e6
        m[60:70, 140:] = (40.0, 0.0)  # off the frame
        d["motion"] = m.astype(np.float32)
    op = O.canonical_params(5, compensated=fp16)
    ref = oracle_canonical_sequence(inputs, op)
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    with S.Context(S.MODE_CANONICAL, storage, max_width=w, max_height=h) as ctx:
        got = gpu_canonical_sequence(ctx, inputs, dt)
    for f, (g, r) in enumerate(zip(got, ref)):
        dm, dmean = rel_err(g["denoised"][..., :3], r["denoised"][..., :3])
        print(f"{form} frame {f}: denoised max {dm:.3e} mean {dmean:.3e}")
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        assert dm <= (1e-3 if fp16 else 2e-4), f


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
def test_canonical_zero_motion(S, oracle_lib, form):
    """A motion plane of zeros is same-pixel history (the reference's temporal rule): 4 frames of
    a static camera with 10 % forced disocclusion vs the FP64 oracle, history lengths exact, and
    every pixel that was never disoccluded has length = frame + 1."""
    O = oracle_lib
    w, h = 130, 70
    fp16 = form != "fp32"
    inputs = seq_inputs(w, h, 4, seed0=90, noise=1 if fp16 else 0, pan_px=0.0, disocclusion=0.1, half=fp16)
    for d in inputs:
        d["motion"] = np.zeros_like(d["motion"])
    ref = oracle_canonical_sequence(inputs, O.canonical_params(5, compensated=fp16))
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    with S.Context(S.MODE_CANONICAL, storage, max_width=w, max_height=h) as ctx:
        got = gpu_canonical_sequence(ctx, inputs, dt)
    for f, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        assert g["history"][..., 3].max() == f + 1, f
        assert rel_err(g["denoised"][..., :3], r["denoised"][..., :3])[0] <= (1e-3 if fp16 else 2e-4), f


@pytest.mark.parametrize("phis", [(4.0, 0.0, 0.0), (4.0, 8.0, 0.5), (2.0, 128.0, 0.0), (10.0, 0.0, 3.0)],
                         ids=["luma_only", "weak_normal", "no_depth", "no_normal"])
def test_canonical_edge_stop_settings(S, oracle_lib, phis):
    """Edge-stop weights other than the defaults, including each of the normal and depth terms
    switched off (phi 0): FP32 GPU vs FP64 oracle over 3 frames."""
    O = oracle_lib
    w, h = 120, 80
    inputs = seq_inputs(w, h, 3, pan_px=1.37, disocclusion=0.1)
    op = O.canonical_params(5)
    gp = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP32)
    for p in (op, gp):
        p.phi_luma, p.phi_normal, p.phi_depth = phis
    ref = oracle_canonical_sequence(inputs, op)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP32, params=gp, max_width=w, max_height=h) as ctx:
        got = gpu_canonical_sequence(ctx, inputs, torch.float32)
    for f, (g, r) in enumerate(zip(got, ref)):
        dm, dmean = rel_err(g["denoised"][..., :3], r["denoised"][..., :3])
        print(f"{phis} frame {f}: denoised max {dm:.3e} mean {dmean:.3e}")
        assert dm <= 2e-4 and dmean <= 1e-6, f


@pytest.mark.parametrize("form", ["fp32", "fp16_wide"])
@pytest.mark.parametrize("iterations,cap", [(1, 32), (2, 1), (4, 63), (7, 8), (10, 32)],
                         ids=["it1", "it2_cap1", "it4_cap63", "it7_cap8", "it10"])
def test_canonical_iterations_and_history_caps(S, oracle_lib, form, iterations, cap):
    """A-trous iteration counts other than 5 change which kernels run: with 1 the first
    iteration (tall tiles, young-pixel variance) writes the output; with 2 the last iteration is
    the first residue-pair stage; 7 and 10 reach strides 64 and 512 (lattices of a few points
    or one, most taps clamped). History caps: 1 (no accumulation: alpha 1), 8, and 63, the
    largest the wide state encodes. 5 frames of 150x100 vs the FP64 oracle, history lengths
    exact; FP32 within 2e-4, the FP16 wide form within 1e-3 max relative error."""
    O = oracle_lib
    w, h = 150, 100
    fp16 = form == "fp16_wide"
    inputs = seq_inputs(w, h, 5, seed0=70, noise=1 if fp16 else 0, pan_px=1.37, disocclusion=0.1, half=fp16)
    op = O.canonical_params(iterations, compensated=fp16)
    op.history_cap = cap
    storage, dt = (S.STORAGE_FP16, torch.float16) if fp16 else (S.STORAGE_FP32, torch.float32)
    gp = S.default_params(S.MODE_CANONICAL, storage)
    gp.iterations, gp.history_cap = iterations, cap
    ref = oracle_canonical_sequence(inputs, op)
    bound = 1e-3 if fp16 else 2e-4
    with S.Context(S.MODE_CANONICAL, storage, params=gp, max_width=w, max_height=h) as ctx:
        assert ctx.wide == fp16
        got = gpu_canonical_sequence(ctx, inputs, dt)
    for f, (g, r) in enumerate(zip(got, ref)):
        dm, dmean = rel_err(g["denoised"][..., :3], r["denoised"][..., :3])
        hm, _ = rel_err(g["history"][..., :3], r["history"][..., :3])
        print(f"{form} it={iterations} cap={cap} frame {f}: denoised max {dm:.3e} mean {dmean:.3e}; "
              f"history max {hm:.3e}")
        assert np.isfinite(g["denoised"]).all(), f
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        assert g["history"][..., 3].max() <= cap, f
        assert dm <= bound and hm <= bound, f
        if fp16:
            me = moment_errors(g["moments"], r["moments"])
            assert me["mean"][0] <= bound and me["second"][0] <= bound, f
    if cap > 5:
        assert got[-1]["history"][..., 3].max() == 5  # every frame valid: lengths reach 5


@pytest.mark.parametrize("wide", [True, False], ids=["wide_state", "plain_fp16_state"])
def test_canonical_fp16_compensated_sequence(S, oracle_lib, wide):
    """Precision form: FP16 storage, compensated moments, lognormal noise, 16 frames, vs the
    FP64 oracle. With the wide state (the default) every output stays within 1e-3 MAX relative
    error at every frame: denoised, colour history, moments (mean and raw second moment).
    Without it the binary16 roundings of the fed-back state compound (about u / alpha = 5
    half-ulps), which is why the wide state exists: the mean error stays small but the max is
    only bounded by 4e-3."""
    O = oracle_lib
    w, h = 256, 160
    inputs = seq_inputs(w, h, 16, noise=1, pan_px=1.37, disocclusion=0.1, half=True)
    op = O.canonical_params(5, compensated=True)
    ref = oracle_canonical_sequence(inputs, op, threads=8)
    p = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP16)
    assert p.flags & S.FLAG_WIDE_STATE
    if not wide:
        p.flags = S.FLAG_COMPENSATED
    bound = 1e-3 if wide else 4e-3
    mean_bound = 3e-4 if wide else 1e-3
    worst = {}
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, params=p, max_width=w, max_height=h) as ctx:
        assert ctx.wide == wide and ctx.moment_channels == (4 if wide else 2)
        got = gpu_canonical_sequence(ctx, inputs, torch.float16)
    for f, (g, r) in enumerate(zip(got, ref)):
        errs = {"denoised": rel_err(g["denoised"][..., :3], r["denoised"][..., :3]),
                "history": rel_err(g["history"][..., :3], r["history"][..., :3])}
        me = moment_errors(g["moments"], r["moments"])
        errs["moments_mean"], errs["moments_second"] = me["mean"], me["second"]
        assert np.isfinite(g["denoised"]).all()
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        for n, (mx, mean) in errs.items():
            worst[n] = max(worst.get(n, 0.0), mx)
            # (the binary16 rounding of the output alone averages ~1.8e-4 relative)
            assert mx <= bound and mean <= mean_bound, (f, n, mx, mean)
    print(f"wide={wide}: max rel err over 16 frames", {n: f"{v:.3e}" for n, v in worst.items()})


def _ragged_shapes():
    return [(1, 1), (2, 9), (65, 5), (96, 96), (130, 70), (64, 4)]


def test_precision_form_long_sequence_does_not_drift(S, oracle_lib):
    """The FP16 wide state over 120 frames (192x128, lognormal noise, 1.37 px/frame pan, 10 %
    disocclusion), well past the history cap of 32 where the fed-back state reaches its steady
    state: every frame's denoised output, colour history and moments within 1e-3 MAX relative
    error of the FP64 oracle, and the worst error of frames 100-119 no larger than 1.5x that of
    frames 30-49 (the residual carried in the state keeps binary16 roundings from compounding)."""
    O = oracle_lib
    w, h, frames = 192, 128, 120
    inputs = seq_inputs(w, h, frames, seed0=500, noise=1, pan_px=1.37, disocclusion=0.1, half=True)
    ref = oracle_canonical_sequence(inputs, O.canonical_params(5, compensated=True), threads=8)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=w, max_height=h) as ctx:
        assert ctx.wide
        got = gpu_canonical_sequence(ctx, inputs, torch.float16)
    per_frame = []
    for f, (g, r) in enumerate(zip(got, ref)):
        me = moment_errors(g["moments"], r["moments"])
        errs = [rel_err(g["denoised"][..., :3], r["denoised"][..., :3])[0],
                rel_err(g["history"][..., :3], r["history"][..., :3])[0], me["mean"][0], me["second"][0]]
        assert np.array_equal(g["history"][..., 3], r["history"][..., 3]), f
        assert max(errs) <= 1e-3, (f, errs)
        per_frame.append(max(errs))
    early, late = max(per_frame[30:50]), max(per_frame[100:120])
    print(f"worst rel err: frames 30-49 {early:.3e}, frames 100-119 {late:.3e}, all {max(per_frame):.3e}")
    assert late <= 1.5 * early + 1e-5, (early, late)


@pytest.mark.parametrize("mode", ["reference", "canonical"])
@pytest.mark.parametrize("launch", ["graph", "direct", "persistent_graph", "persistent_direct"])
def test_batch_matches_single(S, mode, launch):
    """Each stream's result is bit-identical alone (single-frame form) and in a batch, for the
    per-stage launches and the persistent launch, replayed from a graph or launched directly.
    5 batches on 3 internal slots: the persistent form reuses a slot's counters."""
    _batch_vs_single(S, mode, launch)


@pytest.mark.parametrize("mode", ["reference", "canonical"])
@pytest.mark.parametrize("launch", ["graph", "persistent_graph"])
@pytest.mark.parametrize("iterations", [1, 10])
def test_batch_matches_single_iteration_limits(S, mode, launch, iterations):
    """The same at the smallest and largest a-trous iteration counts: 2 stages (the first
    a-trous iteration writes the output) and 11 (strides up to 512, wider than every frame)."""
    _batch_vs_single(S, mode, launch, iterations)


def _batch_vs_single(S, mode, launch, iterations=None):
    m = S.MODE_REFERENCE if mode == "reference" else S.MODE_CANONICAL
    st = S.STORAGE_FP16
    params = S.default_params(m, st)
    if iterations is not None:
        params.iterations = iterations
    dt = torch.float16
    dev = torch.device("cuda")
    shapes = _ragged_shapes()
    frames = 5
    inputs = [seq_inputs(w, h, frames, seed0=10 * i, stream=i, pan_px=0.7, disocclusion=0.1) for i, (w, h) in
              enumerate(shapes)]
    maxw = max(w for w, _ in shapes)
    maxh = max(h for _, h in shapes)
    npx = sum(w * h for w, h in shapes)
    flags = (S.BATCH_PERSISTENT if launch.startswith("persistent") else 0) | \
        (S.BATCH_NO_GRAPH if launch.endswith("direct") else 0)
    with S.Context(m, st, params=params, max_width=maxw, max_height=maxh, max_streams=8, max_batch_frames=8,
                   max_batch_pixels=npx, batch_flags=flags) as bctx:
        sids = [bctx.stream_open(w, h) for w, h in shapes]
        batched = [[None] * frames for _ in shapes]
        keep = []
        for f in range(frames):
            items = []
            for i, (w, h) in enumerate(shapes):
                d = inputs[i][f]
                it = {"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
                      "denoised": torch.empty((h, w, 4), dtype=dt, device=dev)}
                if m == S.MODE_CANONICAL:
                    it["gbuffer"] = torch.from_numpy(d["gbuffer"]).to(dev)
                    it["motion"] = torch.from_numpy(d["motion"]).to(dev)
                items.append(it)
            bctx.denoise_batch(items)
            keep.append(items)
            for i, it in enumerate(items):
                batched[i][f] = it["denoised"]
        torch.cuda.synchronize()
    with S.Context(m, st, params=params, max_width=maxw, max_height=maxh) as sctx:
        for i, (w, h) in enumerate(shapes):
            if m == S.MODE_CANONICAL:
                alone = gpu_canonical_sequence(sctx, inputs[i], dt)
                for f in range(frames):
                    assert np.array_equal(alone[f]["denoised"], batched[i][f].float().cpu().numpy()), (i, f)
            else:
                mom = [torch.zeros((h, w, 2), dtype=dt, device=dev) for _ in range(2)]
                for f in range(frames):
                    d = inputs[i][f]
                    rad = torch.from_numpy(d["radiance"]).to(dev).to(dt)
                    den = torch.empty((h, w, 4), dtype=dt, device=dev)
                    sctx.denoise(radiance=rad, denoised=den, moments_curr=mom[f % 2],
                                 moments_prev=mom[(f + 1) % 2] if f else None, history_valid=f > 0)
                    torch.cuda.synchronize()
                    assert torch.equal(den, batched[i][f]), (i, f)


def test_stream_reset_is_cold(S):
    """svgf_stream_reset gives the same result as a freshly opened stream."""
    dev = torch.device("cuda")
    w, h = 70, 33
    inputs = seq_inputs(w, h, 3, pan_px=0.5)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP32, max_width=w, max_height=h, max_streams=2,
                   max_batch_frames=2, max_batch_pixels=2 * w * h) as ctx:
        a = ctx.stream_open(w, h)
        outs = []
        for f in (0, 1, 2):
            if f == 2:
                ctx.stream_reset(a)
            d = inputs[f]
            den = torch.empty((h, w, 4), device=dev)
            ctx.denoise_batch([{"stream_id": a, "radiance": torch.from_numpy(d["radiance"]).to(dev),
                                "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                                "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": den}])
            outs.append(den)
        b = ctx.stream_open(w, h)
        d = inputs[2]
        den_b = torch.empty((h, w, 4), device=dev)
        ctx.denoise_batch([{"stream_id": b, "radiance": torch.from_numpy(d["radiance"]).to(dev),
                            "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                            "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": den_b}])
        torch.cuda.synchronize()
        assert torch.equal(outs[2], den_b)
        assert not torch.equal(outs[1], den_b)


def test_invalid_arguments(S):
    dev = torch.device("cuda")
    w, h = 32, 16
    with S.Context(S.MODE_REFERENCE, S.STORAGE_FP32, max_width=w, max_height=h) as ctx:
        rad = torch.zeros((h, w, 4), device=dev)
        den = torch.zeros((h, w, 4), device=dev)
        mom = torch.zeros((h, w, 2), device=dev)
        with pytest.raises(S.SvgfError) as e:  # wrong storage format
            ctx.denoise(radiance=rad.half(), denoised=den, moments_curr=mom, history_valid=False)
        assert e.value.status == 1
        with pytest.raises(S.SvgfError) as e:  # missing moments_prev with history
            ctx.denoise(radiance=rad, denoised=den, moments_curr=mom, history_valid=True)
        assert e.value.status == 1
        big = torch.zeros((h, 2 * w, 4), device=dev)
        with pytest.raises(S.SvgfError) as e:  # larger than the context
            ctx.denoise(radiance=big, denoised=big, moments_curr=torch.zeros((h, 2 * w, 2), device=dev),
                        history_valid=False)
        assert e.value.status == 6
        with pytest.raises(S.SvgfError) as e:  # batched form not configured
            ctx.denoise_batch([{"stream_id": 0, "radiance": rad, "denoised": den}])
        assert e.value.status in (2, 5)
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=w, max_height=h, max_streams=2,
                   max_batch_frames=2, max_batch_pixels=w * h, hbm_budget_bytes=1) as ctx:
        with pytest.raises(S.SvgfError) as e:
            ctx.stream_open(w, h)
        assert e.value.status == 4


def test_persistent_wait_limit_reports_timeout(S, monkeypatch):
    """The persistent form's bounded dependency wait, forced from the host: with a 0 us limit
    every wait that finds its dependency unfinished gives up at once (the kernel still ends
    normally). svgf_batch_status then reports SVGF_ERR_TIMEOUT for that very batch, and its
    stream restarts with cold history (frames = 0), so the next batch does not reproject the
    wrong state; a batch that completed normally reports SVGF_OK and advances the stream. A
    context with the default limit on the same inputs reports SVGF_OK for every batch. One
    1280x720 frame per batch: its stage i + 1 tickets directly follow its stage i tickets, so
    workgroups do wait (in a batch of many frames a frame's previous stage has long finished
    when its next stage's tickets come)."""
    dev = torch.device("cuda")
    w, h = 1280, 720
    d = S.synth_frame(w, h, seed=5, frame=0)
    rad = torch.from_numpy(d["radiance"]).to(dev).half()
    g = torch.from_numpy(d["gbuffer"]).to(dev)
    mv = torch.from_numpy(d["motion"]).to(dev)

    def run(ctx, calls):
        sid = ctx.stream_open(w, h)
        out = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
        statuses = []
        for k in range(calls):
            before = ctx.stream_state(sid).frames
            ctx.denoise_batch([{"stream_id": sid, "radiance": rad, "gbuffer": g, "motion": mv, "denoised": out}])
            b = ctx.batch_last()
            assert b == k + 1
            st = ctx.batch_status(b, wait=True)
            statuses.append(st)
            frames = ctx.stream_state(sid).frames
            assert frames == (0 if st == 9 else before + 1), (k, st, before, frames)
        assert ctx.batch_status(calls + 5) == 1  # never enqueued
        return statuses

    kw = dict(max_width=w, max_height=h, max_streams=1, max_batch_frames=1, max_batch_pixels=w * h,
              batch_flags=S.BATCH_PERSISTENT)
    monkeypatch.setenv("SVGF_PERSISTENT_WAIT_US", "0")
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, **kw) as ctx:
        st = run(ctx, 6)
    monkeypatch.delenv("SVGF_PERSISTENT_WAIT_US")
    print("statuses with a 0 us wait limit:", st)
    assert set(st) <= {0, 9} and 9 in st
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, **kw) as ctx:
        assert run(ctx, 6) == [0] * 6


@pytest.mark.parametrize("persistent", [False, True], ids=["per_stage", "persistent"])
def test_timed_stages_inside_graph(S, persistent):
    """SVGF_BATCH_TIMED_STAGES puts event-record nodes around the stage kernels of each batch's
    graph: svgf_stage_times returns positive per-stage times averaged over the completed
    batches (one time for the persistent launch), and the outputs are those of the same batches
    without the events."""
    dev = torch.device("cuda")
    shapes = [(96, 96), (320, 180), (640, 480)]
    npx = sum(w * h for w, h in shapes)
    inputs = [seq_inputs(w, h, 4, seed0=20 * i, stream=i, pan_px=0.6, disocclusion=0.05) for i, (w, h) in
              enumerate(shapes)]
    base = S.BATCH_PERSISTENT if persistent else 0
    res = {}
    for flags in (base, base | S.BATCH_TIMED_STAGES):
        with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=640, max_height=480, max_streams=3,
                       max_batch_frames=3, max_batch_pixels=npx, batch_flags=flags) as ctx:
            sids = [ctx.stream_open(w, h) for w, h in shapes]
            outs = []
            for f in range(4):
                items = []
                for i, (w, h) in enumerate(shapes):
                    d = inputs[i][f]
                    items.append({"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                                  "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                                  "motion": torch.from_numpy(d["motion"]).to(dev),
                                  "denoised": torch.empty((h, w, 4), dtype=torch.float16, device=dev)})
                ctx.denoise_batch(items)
                outs.append([it["denoised"] for it in items])
            torch.cuda.synchronize()
            if flags & S.BATCH_TIMED_STAGES:
                t = ctx.stage_times()
                assert len(t) == (1 if persistent else ctx.params.iterations + 1), t
                assert all(0.0 < x < 100.0 for x in t), t
                with pytest.raises(S.SvgfError) as e:  # nothing completed since
                    ctx.stage_times()
                assert e.value.status == 10
            res[flags] = [[o.view(torch.int16).cpu() for o in row] for row in outs]
    a, b = res.values()
    for f in range(4):
        for i in range(len(shapes)):
            assert torch.equal(a[f][i], b[f][i]), (f, i)


@pytest.mark.parametrize("mode", ["reference", "canonical"])
def test_single_frame_calls_on_two_streams(S, mode):
    """Two single-frame calls through one context on two HIP streams, enqueued back to back
    without a host synchronisation, share the context's scratch: each result equals the one the
    same frame gives alone on one stream, bit for bit (the second call waits on the device for
    the first)."""
    dev = torch.device("cuda")
    w, h = 1280, 720
    m = S.MODE_REFERENCE if mode == "reference" else S.MODE_CANONICAL
    dt = torch.float16
    frames = [S.synth_frame(w, h, seed=30 + k, frame=k, pan_px=0.5, disocclusion=0.05) for k in range(2)]
    with S.Context(m, S.STORAGE_FP16, max_width=w, max_height=h) as ctx:
        def planes(d):
            return {"radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
                    "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev) if m else None,
                    "history_curr": torch.empty((h, w, 4), dtype=dt, device=dev) if m else None,
                    "moments_curr": torch.empty((h, w, ctx.moment_channels), dtype=dt, device=dev),
                    "denoised": torch.empty((h, w, 4), dtype=dt, device=dev), "history_valid": False}
        alone = []
        for d in frames:
            pl = planes(d)
            ctx.denoise(**pl)
            torch.cuda.synchronize()
            alone.append(pl)
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        both = [planes(d) for d in frames]
        torch.cuda.synchronize()
        ctx.denoise(stream=s1, **both[0])
        ctx.denoise(stream=s2, **both[1])
        torch.cuda.synchronize()
        for a, b in zip(alone, both):
            for k in ("denoised", "moments_curr", "history_curr"):
                if a[k] is not None:
                    assert torch.equal(a[k].view(torch.int16), b[k].view(torch.int16)), k


@pytest.mark.parametrize("launch", ["graph", "direct"])
def test_batches_of_one_stream_on_alternating_hip_streams(S, launch):
    """Consecutive batches naming the same agent streams, enqueued back to back on two HIP
    streams in turn without a host synchronisation: each batch reads the state the previous one
    writes, so the library orders them on the device. Every output equals the same sequence on
    one HIP stream, bit for bit."""
    dev = torch.device("cuda")
    dt = torch.float16
    shapes = [(1280, 720), (640, 480), (96, 96)]
    frames = 6
    inputs = [seq_inputs(w, h, frames, seed0=40 + 10 * i, stream=i, pan_px=0.6, disocclusion=0.05, half=True)
              for i, (w, h) in enumerate(shapes)]
    npx = sum(w * h for w, h in shapes)
    flags = S.BATCH_NO_GRAPH if launch == "direct" else 0

    def run(streams):
        with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=1280, max_height=720, max_streams=4,
                       max_batch_frames=4, max_batch_pixels=npx, batch_flags=flags) as ctx:
            sids = [ctx.stream_open(w, h) for w, h in shapes]
            items_all = []
            for f in range(frames):
                items = []
                for i, (w, h) in enumerate(shapes):
                    d = inputs[i][f]
                    items.append({"stream_id": sids[i], "radiance": torch.from_numpy(d["radiance"]).to(dev).to(dt),
                                  "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                                  "motion": torch.from_numpy(d["motion"]).to(dev),
                                  "denoised": torch.empty((h, w, 4), dtype=dt, device=dev)})
                items_all.append(items)
            torch.cuda.synchronize()
            for f in range(frames):
                ctx.denoise_batch(items_all[f], stream=streams[f % len(streams)])
            torch.cuda.synchronize()
            return [[it["denoised"].view(torch.int16).cpu() for it in items] for items in items_all]

    one = run([torch.cuda.Stream()])
    two = run([torch.cuda.Stream(), torch.cuda.Stream()])
    for f in range(frames):
        for i in range(len(shapes)):
            assert torch.equal(one[f][i], two[f][i]), (f, shapes[i])


def test_set_params_keeps_the_state_form_of_open_streams(S):
    """The moment / history form (compensated, wide) cannot change while resident streams hold
    state in the old one; other parameters can. With no stream open the form changes, and the
    moments planes follow it."""
    dev = torch.device("cuda")
    w, h = 64, 32
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=w, max_height=h, max_streams=2, max_batch_frames=2,
                   max_batch_pixels=2 * w * h) as ctx:
        assert ctx.wide and ctx.moment_channels == 4
        sid = ctx.stream_open(w, h)
        plain = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP16)
        plain.flags = S.FLAG_COMPENSATED
        with pytest.raises(S.SvgfError) as e:
            ctx.set_params(plain)
        assert e.value.status == 1 and ctx.wide
        q = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP16)
        q.phi_luma = 6.0
        ctx.set_params(q)
        ctx.stream_close(sid)
        ctx.set_params(plain)
        assert not ctx.wide and ctx.moment_channels == 2
        d = S.synth_frame(w, h, seed=2, frame=0)
        sid = ctx.stream_open(w, h)
        den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
        ctx.denoise_batch([{"stream_id": sid, "radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                            "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
                            "motion": torch.from_numpy(d["motion"]).to(dev), "denoised": den}])
        assert ctx.batch_status(ctx.batch_last()) == 0
        st = ctx.stream_state(sid)
        assert st.moments.format == S.FMT_RG16F


def test_batch_outcomes_outlive_their_slots(S):
    """A batch's outcome stays readable after later batches have reused its slot (a serving
    loop's lane may read it only after the other lanes have cycled): ten batches in a row on one
    stream, then every outcome, oldest first; an outcome 64 batches old is gone."""
    dev = torch.device("cuda")
    w, h = 64, 48
    d = S.synth_frame(w, h, seed=9, frame=0, pan_px=0.0)
    item = {"radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
            "gbuffer": torch.from_numpy(d["gbuffer"]).to(dev),
            "motion": torch.from_numpy(d["motion"]).to(dev),
            "denoised": torch.empty((h, w, 4), dtype=torch.float16, device=dev)}
    with S.Context(S.MODE_CANONICAL, S.STORAGE_FP16, max_width=w, max_height=h, max_streams=1,
                   max_batch_frames=1, max_batch_pixels=w * h) as ctx:
        sid = ctx.stream_open(w, h)
        first = None
        for _ in range(10):
            ctx.denoise_batch([{"stream_id": sid, **item}])
            first = first or ctx.batch_last()
        torch.cuda.synchronize()
        for b in range(first, first + 10):
            assert ctx.batch_status(b, wait=True) == 0, b
        for _ in range(64):
            ctx.denoise_batch([{"stream_id": sid, **item}])
        torch.cuda.synchronize()
        assert ctx.batch_status(first, wait=True) == 1  # evicted: invalid argument
        assert ctx.batch_status(ctx.batch_last(), wait=True) == 0


def test_batch_stage_tile_limit(S):
    """Without the persistent form each stage is one launch of one block per tile, and a launch
    holds at most 2^32 - 1 work-items (16777215 blocks of 256 threads). At 10 passes a REFERENCE
    frame has 262144 single-residue tiles at stride 512 whatever its size: a batch of 64 frames
    is refused with SVGF_ERR_CAPACITY before anything is enqueued (no batch number, every
    stream still at frame 0), one of 63 runs, and each stream's output equals the single-frame
    form's bit for bit. The same 64 frames in CANONICAL mode run in one batch: a lattice of at
    most 40 points runs direct tiles (one per 256 pixels), so a tiny frame has one tile at every
    stride."""
    dev = torch.device("cuda")
    n, w, h = 64, 2, 1
    frames = [S.synth_frame(w, h, seed=300 + i, frame=0) for i in range(n)]
    for mode in (S.MODE_REFERENCE, S.MODE_CANONICAL):
        p = S.default_params(mode, S.STORAGE_FP16)
        p.iterations = 10
        canonical = mode == S.MODE_CANONICAL
        with S.Context(mode, S.STORAGE_FP16, params=p, max_width=w, max_height=h, max_streams=n,
                       max_batch_frames=n, max_batch_pixels=n * w * h) as ctx:
            sids = [ctx.stream_open(w, h) for _ in range(n)]
            items = []
            for sid, d in zip(sids, frames):
                it = {"stream_id": sid, "radiance": torch.from_numpy(d["radiance"]).to(dev).half(),
                      "denoised": torch.empty((h, w, 4), dtype=torch.float16, device=dev)}
                if canonical:
                    it["gbuffer"] = torch.from_numpy(d["gbuffer"]).to(dev)
                    it["motion"] = torch.from_numpy(d["motion"]).to(dev)
                items.append(it)
            run = items
            if not canonical:
                with pytest.raises(S.SvgfError) as e:
                    ctx.denoise_batch(items)
                assert e.value.status == 6  # SVGF_ERR_CAPACITY
                assert ctx.batch_last() == 0
                assert all(ctx.stream_state(sid).frames == 0 for sid in sids)
                run = items[:n - 1]
            ctx.denoise_batch(run)
            assert ctx.batch_status(ctx.batch_last(), wait=True) == 0
            assert all(ctx.stream_state(sid).frames == (1 if i < len(run) else 0) for i, sid in enumerate(sids))
            torch.cuda.synchronize()
        with S.Context(mode, S.STORAGE_FP16, params=p, max_width=w, max_height=h) as sctx:
            for i in (0, 1, 31, len(run) - 1):
                if canonical:
                    alone = gpu_canonical_sequence(sctx, [frames[i]], torch.float16)
                    assert np.array_equal(alone[0]["denoised"], items[i]["denoised"].float().cpu().numpy()), i
                else:
                    mom = torch.zeros((h, w, 2), dtype=torch.float16, device=dev)
                    den = torch.empty((h, w, 4), dtype=torch.float16, device=dev)
                    sctx.denoise(radiance=items[i]["radiance"], denoised=den, moments_curr=mom, history_valid=False)
                    torch.cuda.synchronize()
                    assert torch.equal(den, items[i]["denoised"]), i
