"""CPU-side checks of the C ABI library: it loads, exports every function include/svgf.h
declares, and its host-only functions (parameters, status strings, sizing, synthetic inputs)
behave. No kernel is launched here."""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np
import pytest

import svgf_amd as S


def test_library_exports_every_header_function():
    names = S.header_functions()
    assert len(names) >= 15
    L = S.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(S.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported
    # nothing but the C ABI is exported
    assert {e for e in exported if e.startswith("svgf_")} == set(names)


def test_abi_version_and_status_strings():
    L = S.lib()
    assert L.svgf_abi_version() == 1
    for code, text in S.STATUS.items():
        assert L.svgf_status_str(code).decode() == text
    assert L.svgf_status_str(1234).decode() == "unknown status"


def test_default_params_reference_constants():
    """svgf_denoise.slang:118-150, main.cpp:929."""
    p = S.default_params(S.MODE_REFERENCE)
    assert p.iterations == 3
    assert p.phi_luma == 2.0
    assert p.alpha_moments == pytest.approx(0.2)
    assert p.variance_eps == pytest.approx(1e-4)
    assert p.flags == 0
    c16 = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP16)
    assert c16.iterations == 5 and c16.flags == S.FLAG_COMPENSATED | S.FLAG_WIDE_STATE
    assert S.default_params(S.MODE_CANONICAL, S.STORAGE_FP32).flags == 0


def test_default_params_rejects_bad_input():
    p = S.Params()
    assert S.lib().svgf_default_params(7, 0, ctypes.byref(p)) == 1
    assert S.lib().svgf_default_params(0, 0, None) == 1


def test_query_bytes():
    d = S.Desc(S.MODE_CANONICAL, S.STORAGE_FP16, S.default_params(1, 1), 1280, 720, 4, 8, 4_000_000, 0, 0, 0)
    ctx_b, st_b = ctypes.c_uint64(), ctypes.c_uint64()
    assert S.lib().svgf_query_bytes(ctypes.byref(d), 1280, 720, ctypes.byref(ctx_b), ctypes.byref(st_b)) == 0
    px = 1280 * 720
    # default FP16 form (wide state): history 2 x rgba16f + moments 2 x rgba16f (hi / lo words)
    # + G copy 2 x rgba32f
    assert st_b.value >= px * (2 * 8 + 2 * 8 + 2 * 16)
    assert st_b.value < px * (2 * 8 + 2 * 8 + 2 * 16) + 6 * 256
    # without the wide state: moments 2 x rg16f
    d.params.flags = S.FLAG_COMPENSATED
    assert S.lib().svgf_query_bytes(ctypes.byref(d), 1280, 720, None, ctypes.byref(st_b)) == 0
    assert px * (2 * 8 + 2 * 4 + 2 * 16) <= st_b.value < px * (2 * 8 + 2 * 4 + 2 * 16) + 6 * 256
    d.mode = S.MODE_REFERENCE
    d.storage = S.STORAGE_FP32
    assert S.lib().svgf_query_bytes(ctypes.byref(d), 640, 360, None, ctypes.byref(st_b)) == 0
    assert st_b.value >= 640 * 360 * 2 * 8


def test_create_without_gpu_fails_cleanly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    d = S.Desc(S.MODE_REFERENCE, S.STORAGE_FP32, S.default_params(0), 64, 64, 0, 0, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = S.lib().svgf_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == 8 and not h.value


def test_create_rejects_bad_descriptors():
    h = ctypes.c_void_p()
    L = S.lib()
    bad = S.Desc(S.MODE_REFERENCE, S.STORAGE_FP32, S.default_params(0), 0, 64, 0, 0, 0, 0, 0, 0)
    assert L.svgf_create(ctypes.byref(bad), ctypes.byref(h)) == 1
    bad = S.Desc(S.MODE_REFERENCE, S.STORAGE_FP32, S.default_params(0), 64, 64, 0, 0, 0, 0, 0, 0x80)
    assert L.svgf_create(ctypes.byref(bad), ctypes.byref(h)) == 1
    # every declared batch flag passes validation (here it then stops at the device check)
    known = S.BATCH_NO_GRAPH | S.PROFILE_STAGES | S.BATCH_PERSISTENT | S.BATCH_TIMED_STAGES
    ok = S.Desc(S.MODE_CANONICAL, S.STORAGE_FP16, S.default_params(1, 1), 64, 64, 4, 4, 4096, 0, 0, known)
    try:
        import torch

        gpu = torch.cuda.is_available()
    except ImportError:
        gpu = False
    rc = L.svgf_create(ctypes.byref(ok), ctypes.byref(h))
    if gpu:
        assert rc == 0 and L.svgf_destroy(h) == 0
    else:
        assert rc == 8
    p = S.default_params(0)
    p.iterations = 0
    bad = S.Desc(S.MODE_REFERENCE, S.STORAGE_FP32, p, 64, 64, 0, 0, 0, 0, 0, 0)
    assert L.svgf_create(ctypes.byref(bad), ctypes.byref(h)) == 1
    p = S.default_params(0)
    p.flags = S.FLAG_COMPENSATED  # compensated moments are not the reference's semantics
    bad = S.Desc(S.MODE_REFERENCE, S.STORAGE_FP32, p, 64, 64, 0, 0, 0, 0, 0, 0)
    assert L.svgf_create(ctypes.byref(bad), ctypes.byref(h)) == 2
    # the wide state: FP16 CANONICAL with compensated moments only, lengths up to 63
    for storage, flags, cap, want in ((S.STORAGE_FP32, S.FLAG_COMPENSATED | S.FLAG_WIDE_STATE, 32, 2),
                                      (S.STORAGE_FP16, S.FLAG_WIDE_STATE, 32, 2),
                                      (S.STORAGE_FP16, S.FLAG_COMPENSATED | S.FLAG_WIDE_STATE, 64, 1)):
        p = S.default_params(S.MODE_CANONICAL, storage)
        p.flags, p.history_cap = flags, cap
        bad = S.Desc(S.MODE_CANONICAL, storage, p, 64, 64, 0, 0, 0, 0, 0, 0)
        assert L.svgf_create(ctypes.byref(bad), ctypes.byref(h)) == want, (storage, flags, cap)
    assert L.svgf_create(None, ctypes.byref(h)) == 1
    assert L.svgf_destroy(None) == 1


@pytest.mark.parametrize("field,value", [
    ("iterations", 0), ("iterations", 11), ("phi_luma", 0.0), ("phi_luma", float("nan")), ("phi_luma", -1.0),
    ("phi_normal", -1.0), ("phi_normal", float("nan")), ("phi_normal", float("inf")), ("phi_depth", -0.5),
    ("alpha_color", -0.1), ("alpha_color", 1.5), ("alpha_moments", 1.01), ("alpha_moments", float("nan")),
    ("variance_eps", 0.0), ("variance_eps", -1e-4), ("depth_eps", 0.0), ("depth_reject", float("nan")),
    ("normal_reject", 1.5), ("normal_reject", float("nan")), ("history_cap", 0), ("history_cap", 256),
    ("flags", 0x4)])
def test_create_rejects_each_bad_canonical_parameter(field, value):
    """Every CANONICAL parameter outside its domain is SVGF_ERR_INVALID_ARGUMENT before any
    device is touched (FP32 storage, so the history cap may go up to 255)."""
    p = S.default_params(S.MODE_CANONICAL, S.STORAGE_FP32)
    setattr(p, field, value)
    d = S.Desc(S.MODE_CANONICAL, S.STORAGE_FP32, p, 64, 64, 0, 0, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    assert S.lib().svgf_create(ctypes.byref(d), ctypes.byref(h)) == 1 and not h.value


def test_null_context_calls_fail_cleanly():
    L = S.lib()
    sid = ctypes.c_uint32()
    assert L.svgf_denoise(None, None, None) == 1
    assert L.svgf_stream_open(None, 8, 8, ctypes.byref(sid)) == 1
    assert L.svgf_stream_close(None, 0) == 1
    assert L.svgf_denoise_batch(None, None, 0, None) == 1
    n = ctypes.c_uint32()
    assert L.svgf_stage_times(None, None, 0, ctypes.byref(n)) == 1


# ------------------------------------------------------------------------------------------
# synthetic inputs
# ------------------------------------------------------------------------------------------
def test_synth_is_deterministic_and_thread_independent():
    a = S.synth_frame(97, 41, seed=3, stream=2, frame=5, noise=1, pan_px=1.3, disocclusion=0.1, threads=1)
    b = S.synth_frame(97, 41, seed=3, stream=2, frame=5, noise=1, pan_px=1.3, disocclusion=0.1, threads=7)
    for k in a:
        assert np.array_equal(a[k], b[k])
    c = S.synth_frame(97, 41, seed=4, stream=2, frame=5, noise=1, pan_px=1.3, disocclusion=0.1)
    assert not np.array_equal(a["radiance"], c["radiance"])
    # the scene (normals, depth) does not depend on the noise seed outside forced disocclusions
    e = S.synth_frame(97, 41, seed=4, stream=2, frame=5, noise=1, pan_px=1.3)
    f = S.synth_frame(97, 41, seed=3, stream=2, frame=5, noise=1, pan_px=1.3)
    assert np.array_equal(e["gbuffer"], f["gbuffer"])


def test_synth_ranges():
    d = S.synth_frame(640, 360, seed=0, frame=0, noise=0, pan_px=0.5)
    g = d["gbuffer"]
    assert g[..., 3].min() >= 0.1 - 1e-6 and g[..., 3].max() <= 100.0 + 1e-4
    np.testing.assert_allclose(np.linalg.norm(g[..., :3], axis=-1), 1.0, rtol=1e-6)
    assert np.all(d["motion"][..., 0] == 0.5) and np.all(d["motion"][..., 1] == 0.0)
    rad = d["radiance"]
    assert rad[..., :3].min() >= 0.0 and np.all(rad[..., 3] == 1.0)
    # exponential noise with mean 1 over a 0-10 gradient: mean radiance ~ mean gradient
    base = S.synth_frame(640, 360, seed=0, frame=0, noise=0, pan_px=0.5)["radiance"]
    assert 2.0 < base[..., 0].mean() < 8.0


def test_synth_lognormal_mean_one_and_fp16_range():
    d = S.synth_frame(512, 512, seed=1, frame=0, noise=1)
    r = d["radiance"][..., 0].astype(np.float64)
    d0 = S.synth_frame(512, 512, seed=1, frame=0, noise=0)
    ratio = r.sum() / d0["radiance"][..., 0].astype(np.float64).sum()
    assert 0.85 < ratio < 1.15  # both noises have mean 1
    assert r.max() < 65504.0


def test_synth_noise_free_mode_is_the_mean_of_the_noisy_modes():
    """noise = 2 gives the noise-free radiance: independent of the noise seed, and the mean the
    two noisy modes scatter around (both noises have mean 1)."""
    a = S.synth_frame(256, 256, seed=1, frame=2, noise=2, pan_px=0.5)["radiance"]
    b = S.synth_frame(256, 256, seed=9, frame=2, noise=2, pan_px=0.5)["radiance"]
    assert np.array_equal(a, b)
    for noise in (0, 1):
        n = S.synth_frame(256, 256, seed=1, frame=2, noise=noise, pan_px=0.5)["radiance"]
        ratio = n[..., :3].astype(np.float64).sum() / a[..., :3].astype(np.float64).sum()
        assert 0.9 < ratio < 1.1, (noise, ratio)
    d = S.SynthDesc(8, 8, 0, 0, 0, 3, 0.5, 0.0, 1, 0)
    assert S.lib().svgf_synth_frame(ctypes.byref(d), None, None, None) == 1


def test_synth_pan_motion_is_consistent():
    """Frame f at x shows what frame f-1 shows at x + pan (integer pan: exact)."""
    a = S.synth_frame(64, 16, seed=0, frame=3, pan_px=2.0)
    b = S.synth_frame(64, 16, seed=0, frame=4, pan_px=2.0)
    np.testing.assert_array_equal(b["gbuffer"][:, :-2], a["gbuffer"][:, 2:])


def test_synth_disocclusion_fraction():
    d = S.synth_frame(400, 300, seed=0, frame=2, pan_px=0.0, disocclusion=0.1)
    base = S.synth_frame(400, 300, seed=0, frame=2, pan_px=0.0, disocclusion=0.0)
    frac = np.mean(d["gbuffer"][..., 3] != base["gbuffer"][..., 3])
    assert 0.09 < frac < 0.11


def test_binding_flags_match_header():
    """The ctypes binding's flag constants are the values include/svgf.h defines."""
    import re

    text = (pathlib.Path(__file__).resolve().parent.parent / "include" / "svgf.h").read_text()
    defs = {m.group(1): int(m.group(2), 16) for m in re.finditer(r"#define SVGF_(\w+) (0x[0-9a-fA-F]+)u", text)}
    assert defs["BATCH_NO_GRAPH"] == S.BATCH_NO_GRAPH
    assert defs["PROFILE_STAGES"] == S.PROFILE_STAGES
    assert defs["BATCH_PERSISTENT"] == S.BATCH_PERSISTENT
    assert defs["FLAG_COMPENSATED_MOMENTS"] == S.FLAG_COMPENSATED
    assert defs["FLAG_WIDE_STATE"] == S.FLAG_WIDE_STATE
    assert defs["BATCH_TIMED_STAGES"] == S.BATCH_TIMED_STAGES
    assert defs["REQUEST_KEEP_GBUFFER"] == S.REQUEST_KEEP_GBUFFER


def test_wire_and_readback_argument_checks_without_gpu():
    """Host-side checks of the wire encoder and readback ring (no kernel launched)."""
    L = S.lib()
    p = S.wire_default_params()
    assert (p.exposure, round(p.depth_near, 6), round(p.depth_far, 6)) == (1.0, 0.1, 100.0)
    assert L.svgf_wire_default_params(None) == 1
    h = ctypes.c_void_p()
    assert L.svgf_wire_create(0, 0, ctypes.byref(h)) == 1       # no frames
    assert L.svgf_wire_create(0, 1 << 20, ctypes.byref(h)) == 1  # over the frame limit
    assert L.svgf_wire_encode(None, None, 0, ctypes.byref(p), None) == 1
    assert L.svgf_readback_create(0, 0, 3, ctypes.byref(h)) == 1
    assert L.svgf_readback_create(0, 64, 1, ctypes.byref(h)) == 1  # a ring needs two slots at least
    assert L.svgf_readback_push(None, None, 8, None, None) == 1
    assert L.svgf_readback_view(None, 1, 1, None, None) == 1
    try:
        import torch

        gpu = torch.cuda.is_available()
    except ImportError:
        gpu = False
    if not gpu:
        assert L.svgf_wire_create(0, 16, ctypes.byref(h)) == 8
        assert L.svgf_readback_create(0, 64, 3, ctypes.byref(h)) == 8


def test_composite_argument_checks_without_gpu():
    """Host-side checks of the consumer composite (no kernel launched)."""
    L = S.lib()
    p = S.composite_default_params()
    assert abs(p.ambient - 0.03) < 1e-9  # AMBIENT, mesh.slang:276
    assert L.svgf_composite_default_params(None) == 1
    assert ctypes.sizeof(S.CompositeFrame) == 8 + 8 + 64 + 7 * ctypes.sizeof(S.Plane)
    h = ctypes.c_void_p()
    assert L.svgf_composite_create(0, 0, ctypes.byref(h)) == 1
    assert L.svgf_composite_create(0, 1 << 20, ctypes.byref(h)) == 1
    assert L.svgf_composite_run(None, None, 0, ctypes.byref(p), None) == 1
    assert L.svgf_composite_destroy(None) == 1
    try:
        import torch

        gpu = torch.cuda.is_available()
    except ImportError:
        gpu = False
    if not gpu:
        assert L.svgf_composite_create(0, 16, ctypes.byref(h)) == 8


def test_pack_argument_checks_without_gpu():
    """Host-side checks of the producer pack (no kernel launched)."""
    L = S.lib()
    assert ctypes.sizeof(S.PackFrame) == 16 + 128 + 8 * ctypes.sizeof(S.Plane)
    h = ctypes.c_void_p()
    assert L.svgf_pack_create(0, 0, ctypes.byref(h)) == 1
    assert L.svgf_pack_create(0, 1 << 20, ctypes.byref(h)) == 1
    assert L.svgf_pack_run(None, None, 0, None) == 1
    assert L.svgf_pack_destroy(None) == 1
    try:
        import torch

        gpu = torch.cuda.is_available()
    except ImportError:
        gpu = False
    if not gpu:
        assert L.svgf_pack_create(0, 16, ctypes.byref(h)) == 8


def test_profiler_host_mode_works_without_gpu():
    with S.Profiler(["svgf"], device=-1) as p:
        for k in range(3):
            p.begin_frame()
            p.sample(0, 1000.0 * (k + 1))
            p.end_frame()
        assert p.ema()[0] == 1000.0  # frame 0 read back at frame 2 (two slots)
