"""ctypes wrapper of the CPU oracle (oracle/svgf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product path never loads it.

Arrays are numpy, tightly packed, row-major, channels interleaved: colour-like planes are
(h, w, 4), moments and motion (h, w, 2).
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# SVGF_ORACLE_LIB names another build of the same sources (the sanitizer test's)
LIB_PATH = pathlib.Path(os.environ["SVGF_ORACLE_LIB"]) if "SVGF_ORACLE_LIB" in os.environ else \
    HERE / "_build" / "libsvgf_oracle.so"

FLAG_COMPENSATED = 0x1
FLAG_AS_REFERENCE = 0x100
FLAG_PREFILTER_UNIT = 0x200
FLAG_EPS_OUTSIDE = 0x400
STORE_F32 = 0
STORE_F16 = 1


class OracleParams(ctypes.Structure):
    """Field order of svgf_params (include/svgf.h) and oracle_params."""

    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("phi_luma", ctypes.c_float),
        ("phi_normal", ctypes.c_float),
        ("phi_depth", ctypes.c_float),
        ("alpha_color", ctypes.c_float),
        ("alpha_moments", ctypes.c_float),
        ("variance_eps", ctypes.c_float),
        ("depth_eps", ctypes.c_float),
        ("depth_reject", ctypes.c_float),
        ("normal_reject", ctypes.c_float),
        ("history_cap", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


def reference_params(iterations: int = 3) -> OracleParams:
    """The shader's constants: svgf_denoise.slang:118-150, strides main.cpp:929."""
    return OracleParams(iterations, 2.0, 0.0, 0.0, 1.0, 0.2, 1e-4, 0.0, 0.0, 0.0, 1, 0)


def canonical_params(iterations: int = 5, compensated: bool = False) -> OracleParams:
    """Canonical SVGF defaults (Schied et al. 2017): phi_l 4, phi_n 128, phi_z 1."""
    return OracleParams(iterations, 4.0, 128.0, 1.0, 0.2, 0.2, 1e-4, 1e-3, 0.1, 0.9, 32,
                        FLAG_COMPENSATED if compensated else 0)


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OracleParams)
        L.oracle_r_frame_f32.argtypes = [ctypes.c_int, ctypes.c_int, pp, ctypes.c_int, ctypes.c_int,
                                         fp, fp, fp, fp, fp, fp, ctypes.c_int]
        L.oracle_r_frame_f64.argtypes = [ctypes.c_int, ctypes.c_int, pp, ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.oracle_c_frame_f64.argtypes = [ctypes.c_int, ctypes.c_int, pp, ctypes.c_int, fp, fp, fp, fp,
                                         dp, dp, dp, dp, dp, dp, dp]
        L.oracle_c_frame_f64_mt.argtypes = L.oracle_c_frame_f64.argtypes + [ctypes.c_int]
        L.oracle_c_frame_f32_mt.argtypes = [ctypes.c_int, ctypes.c_int, pp, ctypes.c_int, fp, fp, fp, fp,
                                            fp, fp, fp, fp, fp, fp, fp, ctypes.c_int]
        L.oracle_f32_to_f16.argtypes = [ctypes.c_float]
        L.oracle_f32_to_f16.restype = ctypes.c_uint16
        L.oracle_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.oracle_f16_to_f32.restype = ctypes.c_float
        L.oracle_round_f16_array.argtypes = [fp, ctypes.c_int64]
        L.oracle_wire_thresholds.argtypes = [fp]
        L.oracle_wire_tonemap.argtypes = [fp, ctypes.c_float, fp]
        L.oracle_wire_codes.argtypes = [fp, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8)]
        L.oracle_wire_depth.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.oracle_wire_depth.restype = ctypes.c_float
        L.oracle_wire_encode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         fp, fp, ctypes.POINTER(ctypes.c_uint8), fp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_composite_lookup.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_float, ip, ip]
        L.oracle_composite.argtypes = [ctypes.c_int, ctypes.c_int, fp, ctypes.c_float, fp, fp, fp, fp, fp, fp,
                                       ctypes.c_int, fp]
        up = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_unpack_octahedral.argtypes = [ctypes.c_uint32, fp]
        L.oracle_profiler_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ctypes.c_float, ctypes.c_uint32,
                                          ctypes.c_uint32, fp, ctypes.POINTER(ctypes.c_uint8)]
        L.oracle_pack.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp, fp, up, fp, fp, fp, ctypes.c_int, fp, fp, fp]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def round_f16(a: np.ndarray) -> np.ndarray:
    """binary16 round-to-nearest-even of a float32 array (the oracle's own conversion)."""
    out = np.array(a, dtype=np.float32, copy=True, order="C")
    lib().oracle_round_f16_array(_fp(out), out.size)
    return out


def r_frame_f32(lighting, moments_prev, params=None, store=STORE_F32, history_valid=True, threads=1):
    """One REFERENCE-mode frame with float arithmetic. Returns (moments_curr, denoised,
    wavelet_a, wavelet_b)."""
    p = params or reference_params()
    lighting = _f32(lighting)
    h, w = lighting.shape[:2]
    mp = _f32(moments_prev) if moments_prev is not None else np.zeros((h, w, 2), np.float32)
    mc = np.zeros((h, w, 2), np.float32)
    den = np.zeros((h, w, 4), np.float32)
    wa = np.zeros((h, w, 4), np.float32)
    wb = np.zeros((h, w, 4), np.float32)
    rc = lib().oracle_r_frame_f32(w, h, ctypes.byref(p), store, int(bool(history_valid)), _fp(lighting),
                                  _fp(mp), _fp(mc), _fp(den), _fp(wa), _fp(wb), int(threads))
    if rc:
        raise ValueError(f"oracle_r_frame_f32 failed: {rc}")
    return mc, den, wa, wb


def r_frame_f64(lighting, moments_prev, params=None, history_valid=True):
    """One REFERENCE-mode frame in double. Returns (moments_curr, denoised)."""
    p = params or reference_params()
    lighting = _f64(lighting)
    h, w = lighting.shape[:2]
    mp = _f64(moments_prev) if moments_prev is not None else np.zeros((h, w, 2))
    mc = np.zeros((h, w, 2))
    den = np.zeros((h, w, 4))
    wa = np.zeros((h, w, 4))
    wb = np.zeros((h, w, 4))
    rc = lib().oracle_r_frame_f64(w, h, ctypes.byref(p), int(bool(history_valid)), _dp(lighting), _dp(mp),
                                  _dp(mc), _dp(den), _dp(wa), _dp(wb))
    if rc:
        raise ValueError(f"oracle_r_frame_f64 failed: {rc}")
    return mc, den


def c_frame_f64(radiance, gbuffer, gbuffer_prev, motion, history_prev, moments_prev, params=None,
                history_valid=True, debug=False, threads=1):
    """One CANONICAL frame in double. Returns dict with history, moments, denoised (and
    temporal / variance stage outputs when debug)."""
    p = params or canonical_params()
    radiance = _f32(radiance)
    h, w = radiance.shape[:2]
    hist_prev = _f64(history_prev) if history_prev is not None else np.zeros((h, w, 4))
    mom_prev = _f64(moments_prev) if moments_prev is not None else np.zeros((h, w, 2))
    out = {
        "history": np.zeros((h, w, 4)),
        "moments": np.zeros((h, w, 2)),
        "denoised": np.zeros((h, w, 4)),
    }
    if debug:
        out["temporal"] = np.zeros((h, w, 4))
        out["variance"] = np.zeros((h, w, 4))
    gp = _f32(gbuffer_prev) if gbuffer_prev is not None else np.zeros((h, w, 4), np.float32)
    rc = lib().oracle_c_frame_f64_mt(
        w, h, ctypes.byref(p), int(bool(history_valid)), _fp(radiance), _fp(_f32(gbuffer)), _fp(gp),
        _fp(_f32(motion)), _dp(hist_prev), _dp(mom_prev), _dp(out["history"]), _dp(out["moments"]),
        _dp(out["denoised"]), _dp(out.get("temporal")), _dp(out.get("variance")), int(threads))
    if rc:
        raise ValueError(f"oracle_c_frame_f64 failed: {rc}")
    return out


def c_frame_f32(radiance, gbuffer, gbuffer_prev, motion, history_prev, moments_prev, params=None,
                history_valid=True, threads=1):
    """One CANONICAL frame in float: the same code as c_frame_f64 in binary32 (the CPU baseline
    port of the FP32 canonical form, not an oracle). Returns dict with history, moments,
    denoised (float32)."""
    p = params or canonical_params()
    radiance = _f32(radiance)
    h, w = radiance.shape[:2]
    hist_prev = _f32(history_prev) if history_prev is not None else np.zeros((h, w, 4), np.float32)
    mom_prev = _f32(moments_prev) if moments_prev is not None else np.zeros((h, w, 2), np.float32)
    out = {k: np.zeros((h, w, c), np.float32) for k, c in (("history", 4), ("moments", 2), ("denoised", 4))}
    gp = _f32(gbuffer_prev) if gbuffer_prev is not None else np.zeros((h, w, 4), np.float32)
    rc = lib().oracle_c_frame_f32_mt(
        w, h, ctypes.byref(p), int(bool(history_valid)), _fp(radiance), _fp(_f32(gbuffer)), _fp(gp),
        _fp(_f32(motion)), _fp(hist_prev), _fp(mom_prev), _fp(out["history"]), _fp(out["moments"]),
        _fp(out["denoised"]), None, None, int(threads))
    if rc:
        raise ValueError(f"oracle_c_frame_f32 failed: {rc}")
    return out


# ---------------------------------------------------------------------------------------------
# agent-frame wire encoder (oracle/wire_oracle.c)
# ---------------------------------------------------------------------------------------------
WIRE_NEAR, WIRE_FAR = 0.1, 100.0  # camera planes, src/camera.hpp:99-100


def wire_thresholds() -> np.ndarray:
    """The 255 float decision thresholds of the 8-bit sRGB code (ascending)."""
    t = np.zeros(255, np.float32)
    lib().oracle_wire_thresholds(_fp(t))
    return t


def wire_codes(x) -> np.ndarray:
    """8-bit sRGB codes of linear values (any shape)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint8)
    lib().oracle_wire_codes(_fp(x), x.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def wire_tonemap(rgb, exposure: float = 1.0) -> np.ndarray:
    """Fitted ACES of one (3,) colour (float arithmetic), in [0, 1]."""
    src = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(3, np.float32)
    lib().oracle_wire_tonemap(_fp(src), float(exposure), _fp(out))
    return out


def wire_encode(colour, gbuffer, exposure: float = 1.0, near: float = WIRE_NEAR, far: float = WIRE_FAR):
    """colour (h, w, 4) and gbuffer (h, w, 4) float32 -> (rgb (h, w, 3) uint8, depth (h, w) float32)."""
    colour = _f32(colour)
    gbuffer = _f32(gbuffer)
    h, w = colour.shape[:2]
    rgb = np.zeros((h, w, 3), np.uint8)
    depth = np.zeros((h, w), np.float32)
    rc = lib().oracle_wire_encode(w, h, float(exposure), float(near), float(far), _fp(colour), _fp(gbuffer),
                                  rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _fp(depth))
    if rc:
        raise ValueError(f"oracle_wire_encode failed: {rc}")
    return rgb, depth


COMPOSITE_AMBIENT = 0.03  # src/mesh.slang:276
KIND_NONE, KIND_READY, KIND_REPROJECT = 0.0, 1.0, 2.0


def composite_lookup(m, w: int, h: int, pos):
    """History pixel (x, y) a world position reprojects to through the row-major 4x4 m, or None
    (behind the camera or off screen) — mesh.slang:1162-1170."""
    m = _f32(m).reshape(16)
    px, py = ctypes.c_int(0), ctypes.c_int(0)
    rc = lib().oracle_composite_lookup(_fp(m), int(w), int(h), float(pos[0]), float(pos[1]), float(pos[2]),
                                       ctypes.byref(px), ctypes.byref(py))
    return None if rc else (px.value, py.value)


def composite(m, history, emissive, fresnel, reflection, refl_colour, refl_base, ambient: float = COMPOSITE_AMBIENT,
              store=STORE_F32):
    """Every plane (h, w, 4) float32 -> composited colour (h, w, 4) float32 (mesh.slang:1007,
    1161-1174, 1191); store=STORE_F16 rounds the output to half."""
    planes = [_f32(a) for a in (history, emissive, fresnel, reflection, refl_colour, refl_base)]
    h, w = planes[0].shape[:2]
    for a in planes:
        if a.shape != (h, w, 4):
            raise ValueError("every plane is (h, w, 4)")
    out = np.zeros((h, w, 4), np.float32)
    m = _f32(m).reshape(16)
    rc = lib().oracle_composite(w, h, _fp(m), float(ambient), *[_fp(a) for a in planes], int(store), _fp(out))
    if rc:
        raise ValueError(f"oracle_composite failed: {rc}")
    return out


PACK_NO_MOTION = -1.0e6  # motion of a pixel with no previous position


def unpack_octahedral(packed: int) -> np.ndarray:
    """The unit normal of a 16 + 16-bit octahedral code (mesh.slang:341-352)."""
    n = np.zeros(3, np.float32)
    lib().oracle_unpack_octahedral(int(packed) & 0xFFFFFFFF, _fp(n))
    return n


def pack(vp, pvp, position, normal_oct, indirect, direct_diffuse, direct_specular, store=STORE_F32):
    """Producer pack (main.cpp:595-646, mesh.slang:995-998): returns (radiance (h, w, 4),
    gbuffer (h, w, 4), motion (h, w, 2)) float32."""
    position, indirect, direct_diffuse, direct_specular = (_f32(a) for a in (position, indirect, direct_diffuse,
                                                                              direct_specular))
    normal_oct = np.ascontiguousarray(normal_oct, dtype=np.uint32)
    h, w = position.shape[:2]
    for a in (indirect, direct_diffuse, direct_specular):
        if a.shape != (h, w, 4):
            raise ValueError("lighting planes are (h, w, 4)")
    if normal_oct.shape != (h, w):
        raise ValueError("normal_oct is (h, w)")
    rad = np.zeros((h, w, 4), np.float32)
    g = np.zeros((h, w, 4), np.float32)
    mv = np.zeros((h, w, 2), np.float32)
    rc = lib().oracle_pack(w, h, _fp(_f32(vp).reshape(16)), _fp(_f32(pvp).reshape(16)), _fp(position),
                           normal_oct.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _fp(indirect),
                           _fp(direct_diffuse), _fp(direct_specular), int(store), _fp(rad), _fp(g), _fp(mv))
    if rc:
        raise ValueError(f"oracle_pack failed: {rc}")
    return rad, g, mv


def profiler_run(samples, slots=2, alpha=0.05, interval=60, first=16):
    """The reference's profiler loop over per-iteration pass samples (frames, passes) in ns:
    returns (ema (frames, passes) float32 after each iteration's readback, logged (frames,) bool)."""
    samples = _f32(samples)
    frames, passes = samples.shape
    ema = np.zeros((frames, passes), np.float32)
    logged = np.zeros(frames, np.uint8)
    rc = lib().oracle_profiler_run(passes, slots, frames, _fp(samples), float(alpha), interval, first, _fp(ema),
                                   logged.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc:
        raise ValueError(f"oracle_profiler_run failed: {rc}")
    return ema, logged.astype(bool)
