"""Python host binding of the MI355X SVGF library (C ABI of include/svgf.h via ctypes).

The compute path is the HIP library tron-grid_amd/lib/libsvgf_mi355x.so; this module only
marshals descriptors. PyTorch tensors serve as device memory and streams. There is no CPU
fallback: loading fails loudly when the library has not been built.

The single-frame call mirrors TronGrid's SVGF pass (src/main.cpp:907-947 recording the four
dispatches of src/svgf_denoise.slang:152-275): same images, same parameters, same ping-pong of
moments / history, with cold history = zero images (src/main.cpp:1449-1455).
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import re
import subprocess

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent.parent  # tron-grid_amd/
REPO_DIR = PKG_DIR.parent
LIB_PATH = PKG_DIR / "lib" / "libsvgf_mi355x.so"
HEADER_PATH = REPO_DIR / "include" / "svgf.h"

MODE_REFERENCE = 0
MODE_CANONICAL = 1
STORAGE_FP32 = 0
STORAGE_FP16 = 1
FMT_RGBA32F, FMT_RGBA16F, FMT_RG32F, FMT_RG16F, FMT_R32UI = 1, 2, 3, 4, 5
FLAG_COMPENSATED = 0x1
FLAG_WIDE_STATE = 0x2
BATCH_NO_GRAPH = 0x1
PROFILE_STAGES = 0x2
BATCH_PERSISTENT = 0x4
BATCH_TIMED_STAGES = 0x8
REQUEST_KEEP_GBUFFER = 0x1

STATUS = {
    0: "ok", 1: "invalid argument", 2: "unsupported", 3: "out of device memory", 4: "over HBM budget",
    5: "no such stream", 6: "over context capacity", 7: "HIP runtime error", 8: "no gfx950 device", 9: "timeout",
    10: "not ready",
}


class SvgfError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")
        self.status = status


class Params(ctypes.Structure):
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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Desc(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("storage", ctypes.c_uint32),
        ("params", Params),
        ("max_width", ctypes.c_uint32),
        ("max_height", ctypes.c_uint32),
        ("max_streams", ctypes.c_uint32),
        ("max_batch_frames", ctypes.c_uint32),
        ("max_batch_pixels", ctypes.c_uint64),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("device", ctypes.c_int32),
        ("batch_flags", ctypes.c_uint32),
    ]


class Plane(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch_bytes", ctypes.c_uint64), ("format", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class Frame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("history_valid", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)] + [(n, Plane) for n in (
                    "radiance", "gbuffer", "gbuffer_prev", "motion", "history_prev", "history_curr",
                    "moments_prev", "moments_curr", "denoised")]


class Request(ctypes.Structure):
    _fields_ = [("stream_id", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("radiance", Plane),
                ("gbuffer", Plane), ("motion", Plane), ("denoised", Plane)]


class WireParams(ctypes.Structure):
    _fields_ = [("exposure", ctypes.c_float), ("depth_near", ctypes.c_float), ("depth_far", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class WireFrame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("colour", Plane), ("gbuffer", Plane),
                ("rgb", ctypes.c_void_p), ("depth", ctypes.c_void_p)]


COMPOSITE_PLANES = ("history", "emissive", "fresnel", "reflection", "reflection_colour", "reflection_base", "out")


class CompositeFrame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2),
                ("prev_view_projection", ctypes.c_float * 16)] + [(n, Plane) for n in COMPOSITE_PLANES]


PACK_INPUTS = ("position", "normal_oct", "indirect", "direct_diffuse", "direct_specular")
PACK_OUTPUTS = ("radiance", "gbuffer", "motion")


class PackFrame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2),
                ("view_projection", ctypes.c_float * 16), ("prev_view_projection", ctypes.c_float * 16)] + [
                    (n, Plane) for n in PACK_INPUTS + PACK_OUTPUTS]


class ProfilerDesc(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_uint32), ("slots", ctypes.c_uint32), ("ema_alpha", ctypes.c_float),
                ("log_interval", ctypes.c_uint32), ("first_frame", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("names", ctypes.POINTER(ctypes.c_char_p))]


class CompositeParams(ctypes.Structure):
    _fields_ = [("ambient", ctypes.c_float), ("reserved", ctypes.c_uint32 * 3)]


class StreamState(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("frames", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("history", Plane), ("moments", Plane), ("gbuffer", Plane)]


class SynthDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_uint64), ("frame", ctypes.c_uint32), ("noise", ctypes.c_uint32),
        ("pan_px", ctypes.c_float), ("disocclusion", ctypes.c_float), ("threads", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class ServerDesc(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_uint32), ("max_batch_frames", ctypes.c_uint32),
                ("max_batch_pixels", ctypes.c_uint64), ("max_wait_us", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class Completion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("stream_id", ctypes.c_uint32), ("status", ctypes.c_int32),
                ("submit_ns", ctypes.c_uint64), ("launch_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64),
                ("batch_frames", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class SimDesc(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_uint32), ("n_classes", ctypes.c_uint32),
                ("class_width", ctypes.POINTER(ctypes.c_uint32)), ("class_height", ctypes.POINTER(ctypes.c_uint32)),
                ("class_inputs", ctypes.POINTER(Request)), ("inputs_per_class", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("rate_hz", ctypes.c_double), ("duration_s", ctypes.c_double),
                ("churn_per_s", ctypes.c_double), ("seed", ctypes.c_uint64)]


class SimResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "rejected", "batches", "streams_opened",
                                                "streams_closed", "pixels")] + \
               [(n, ctypes.c_double) for n in ("wall_s", "mpx_per_s", "latency_p50_ms", "latency_p99_ms",
                                               "latency_max_ms", "latency_mean_ms", "mean_batch_frames",
                                               "offered_mpx_per_s")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build(force: bool = False) -> None:
    """Compile the HIP library in-tree (hipcc --offload-arch=gfx950)."""
    args = ["make", "-s", "-C", str(PKG_DIR), f"-j{min(8, os.cpu_count() or 1)}"]
    if force:
        subprocess.run(["make", "-s", "-C", str(PKG_DIR), "clean"], check=True)
    subprocess.run(args, check=True)


def header_functions() -> list[str]:
    """Every function include/svgf.h declares."""
    text = HEADER_PATH.read_text()
    return re.findall(r"^SVGF_API\s+[\w\s\*]+?\b(svgf_\w+)\s*\(", text, flags=re.M)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.svgf_abi_version.restype = ctypes.c_uint32
        L.svgf_status_str.restype = ctypes.c_char_p
        L.svgf_status_str.argtypes = [ctypes.c_int]
        for name, args in {
            "svgf_default_params": [ctypes.c_uint32, ctypes.c_uint32, P(Params)],
            "svgf_query_bytes": [P(Desc), ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint64), P(ctypes.c_uint64)],
            "svgf_create": [P(Desc), P(vp)],
            "svgf_destroy": [vp],
            "svgf_set_params": [vp, P(Params)],
            "svgf_denoise": [vp, P(Frame), vp],
            "svgf_denoise_static": [ctypes.c_uint32, ctypes.c_uint32, P(Params), P(Frame), vp],
            "svgf_stream_open": [vp, ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint32)],
            "svgf_stream_close": [vp, ctypes.c_uint32],
            "svgf_stream_reset": [vp, ctypes.c_uint32, vp],
            "svgf_stream_state_get": [vp, ctypes.c_uint32, P(StreamState)],
            "svgf_budget_used": [vp, P(ctypes.c_uint64)],
            "svgf_denoise_batch": [vp, P(Request), ctypes.c_uint32, vp],
            "svgf_batch_last": [vp, P(ctypes.c_uint64)],
            "svgf_batch_status": [vp, ctypes.c_uint64, ctypes.c_int32],
            "svgf_synth_frame": [P(SynthDesc), vp, vp, vp],
            "svgf_stage_times": [vp, P(ctypes.c_float), ctypes.c_uint32, P(ctypes.c_uint32)],
            "svgf_server_create": [vp, P(ServerDesc), P(vp)],
            "svgf_server_destroy": [vp],
            "svgf_server_stream_open": [vp, ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint32)],
            "svgf_server_stream_close": [vp, ctypes.c_uint32],
            "svgf_server_submit": [vp, P(Request), ctypes.c_uint64],
            "svgf_server_poll": [vp, P(Completion), ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32],
            "svgf_server_drain": [vp],
            "svgf_server_simulate": [vp, P(SimDesc), P(SimResult)],
            "svgf_wire_default_params": [P(WireParams)],
            "svgf_wire_create": [ctypes.c_int32, ctypes.c_uint32, P(vp)],
            "svgf_wire_destroy": [vp],
            "svgf_wire_encode": [vp, P(WireFrame), ctypes.c_uint32, P(WireParams), vp],
            "svgf_readback_create": [ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint32, P(vp)],
            "svgf_readback_destroy": [vp],
            "svgf_readback_push": [vp, vp, ctypes.c_uint64, vp, P(ctypes.c_uint64)],
            "svgf_readback_view": [vp, ctypes.c_uint64, ctypes.c_int32, P(vp), P(ctypes.c_uint64)],
            "svgf_composite_default_params": [P(CompositeParams)],
            "svgf_composite_create": [ctypes.c_int32, ctypes.c_uint32, P(vp)],
            "svgf_composite_destroy": [vp],
            "svgf_composite_run": [vp, P(CompositeFrame), ctypes.c_uint32, P(CompositeParams), vp],
            "svgf_pack_create": [ctypes.c_int32, ctypes.c_uint32, P(vp)],
            "svgf_pack_destroy": [vp],
            "svgf_pack_run": [vp, P(PackFrame), ctypes.c_uint32, vp],
            "svgf_profiler_default_desc": [P(ProfilerDesc)],
            "svgf_profiler_create": [P(ProfilerDesc), P(vp)],
            "svgf_profiler_destroy": [vp],
            "svgf_profiler_begin_frame": [vp, P(ctypes.c_int32)],
            "svgf_profiler_pass_begin": [vp, ctypes.c_uint32, vp],
            "svgf_profiler_pass_end": [vp, ctypes.c_uint32, vp],
            "svgf_profiler_sample": [vp, ctypes.c_uint32, ctypes.c_float],
            "svgf_profiler_end_frame": [vp],
            "svgf_profiler_ema": [vp, P(ctypes.c_float), ctypes.c_uint32, P(ctypes.c_uint32)],
            "svgf_profiler_log_line": [vp, ctypes.c_char_p, ctypes.c_uint32],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        raise SvgfError(status, what)


def default_params(mode: int, storage: int = STORAGE_FP32) -> Params:
    p = Params()
    check(lib().svgf_default_params(mode, storage, ctypes.byref(p)), "svgf_default_params")
    return p


# ---------------------------------------------------------------------------------------------
# synthetic inputs (host, no GPU)
# ---------------------------------------------------------------------------------------------
def synth_frame(w, h, *, seed=0, stream=0, frame=0, noise=0, pan_px=0.5, disocclusion=0.0, threads=0,
                radiance=True, gbuffer=True, motion=True):
    """Seeded synthetic inputs: dict of float32 arrays radiance (h,w,4), gbuffer (h,w,4),
    motion (h,w,2)."""
    d = SynthDesc(w, h, seed, stream, frame, noise, pan_px, disocclusion, threads, 0)
    out = {}
    if radiance:
        out["radiance"] = np.empty((h, w, 4), np.float32)
    if gbuffer:
        out["gbuffer"] = np.empty((h, w, 4), np.float32)
    if motion:
        out["motion"] = np.empty((h, w, 2), np.float32)

    def ptr(k):
        return out[k].ctypes.data if k in out else None

    check(lib().svgf_synth_frame(ctypes.byref(d), ptr("radiance"), ptr("gbuffer"), ptr("motion")),
          "svgf_synth_frame")
    return out


# ---------------------------------------------------------------------------------------------
# device planes (torch as allocator)
# ---------------------------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def plane_of(t, fmt: int) -> Plane:
    """Plane descriptor of a contiguous (h, w, c) torch tensor on the GPU."""
    if t is None:
        return Plane()
    assert t.is_cuda and t.is_contiguous(), "planes are contiguous device tensors"
    return Plane(t.data_ptr(), t.stride(0) * t.element_size(), fmt, 0)


def storage_dtype(storage: int):
    torch = _torch()
    return torch.float32 if storage == STORAGE_FP32 else torch.float16


def fmt4(storage):
    return FMT_RGBA32F if storage == STORAGE_FP32 else FMT_RGBA16F


def fmt2(storage):
    return FMT_RG32F if storage == STORAGE_FP32 else FMT_RG16F


def moment_fmt(storage, flags):
    """Format of the moments planes: the wide state's four binary16 words, or two channels."""
    return FMT_RGBA16F if flags & FLAG_WIDE_STATE else fmt2(storage)


# ---------------------------------------------------------------------------------------------
# wide FP16 state (svgf.h SVGF_FLAG_WIDE_STATE; csrc/svgf_state.hpp)
# ---------------------------------------------------------------------------------------------
def _wide_step(h16):
    """2^(max(e, 1) - 28) for binary16 values h16 (e = biased exponent)."""
    e = (np.asarray(h16, np.float16).view(np.uint16).astype(np.int32) >> 10) & 31
    return np.ldexp(1.0, np.maximum(e, 1) - 28)


def decode_history(h):
    """Wide history plane (h, w, 4) float16 -> (colour (h, w, 3) float64, length (h, w) int)."""
    h = np.asarray(h, np.float16)
    word = h[..., 3].view(np.uint16).astype(np.int32)
    rgb = np.empty(h.shape[:-1] + (3,), np.float64)
    for c in range(3):
        q = (word >> (6 + 3 * c)) & 7
        q = np.where(q >= 4, q - 8, q)
        v = h[..., c].astype(np.float64)
        rgb[..., c] = v + np.where(np.isfinite(v), q * _wide_step(h[..., c]), 0.0)
    return rgb, word & 63


def decode_moments(m):
    """Wide moments plane (h, w, 4) float16 (mean_hi, sigma_hi, mean_lo, sigma_lo) ->
    (h, w, 2) float64 (mean, sigma): hi + lo 2^(max(e, 1) - 25), e the biased exponent of hi."""
    m = np.asarray(m, np.float16)
    out = []
    for c in range(2):
        hi = m[..., c]
        lo = m[..., c + 2].astype(np.float64)
        out.append(hi.astype(np.float64) + lo * (_wide_step(hi) * 8.0))
    return np.stack(out, axis=-1)


def make_frame(storage, moment_format, *, radiance, denoised, moments_curr, moments_prev=None, gbuffer=None,
               gbuffer_prev=None, motion=None, history_prev=None, history_curr=None, history_valid=True) -> Frame:
    """svgf_frame of device tensors (h, w, c)."""
    f = Frame()
    h, w = radiance.shape[:2]
    f.width, f.height, f.history_valid = w, h, int(bool(history_valid))
    f.radiance = plane_of(radiance, fmt4(storage))
    f.denoised = plane_of(denoised, fmt4(storage))
    f.moments_curr = plane_of(moments_curr, moment_format)
    f.moments_prev = plane_of(moments_prev, moment_format)
    f.gbuffer = plane_of(gbuffer, FMT_RGBA32F)
    f.gbuffer_prev = plane_of(gbuffer_prev, FMT_RGBA32F)
    f.motion = plane_of(motion, FMT_RG32F)
    f.history_prev = plane_of(history_prev, fmt4(storage))
    f.history_curr = plane_of(history_curr, fmt4(storage))
    return f


STATIC_MAX_PIXELS = 1280 * 720


def denoise_static(mode, storage, params=None, *, stream=None, **planes) -> int:
    """svgf_denoise_static (no context, no allocation) on the current device; returns the
    svgf_status (0 on success) instead of raising, so callers can see each error code."""
    p = params if params is not None else default_params(mode, storage)
    f = make_frame(storage, moment_fmt(storage, p.flags), **planes)
    return lib().svgf_denoise_static(mode, storage, ctypes.byref(p), ctypes.byref(f), _stream_handle(stream))


class Context:
    """One svgf_context (replaces the SVGF pipeline + descriptor sets of src/main.cpp:2288-2446)."""

    def __init__(self, mode=MODE_REFERENCE, storage=STORAGE_FP32, params=None, max_width=1920, max_height=1080,
                 max_streams=0, max_batch_frames=0, max_batch_pixels=0, hbm_budget_bytes=0, device=0,
                 batch_flags=0):
        self.mode, self.storage = mode, storage
        self.params = params if params is not None else default_params(mode, storage)
        self.desc = Desc(mode, storage, self.params, max_width, max_height, max_streams, max_batch_frames,
                         max_batch_pixels, hbm_budget_bytes, device, batch_flags)
        self.moment_format = moment_fmt(storage, self.params.flags)
        self.moment_channels = 4 if self.moment_format == FMT_RGBA16F else 2
        h = ctypes.c_void_p()
        check(lib().svgf_create(ctypes.byref(self.desc), ctypes.byref(h)), "svgf_create")
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().svgf_destroy(self.handle), "svgf_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: Params):
        check(lib().svgf_set_params(self.handle, ctypes.byref(p)), "svgf_set_params")
        self.params = p
        self.moment_format = moment_fmt(self.storage, p.flags)
        self.moment_channels = 4 if self.moment_format == FMT_RGBA16F else 2

    @property
    def wide(self) -> bool:
        return bool(self.params.flags & FLAG_WIDE_STATE)

    def denoise(self, *, stream=None, **planes):
        """Enqueue one frame on `stream` (a torch.cuda.Stream or raw handle; default: current).
        planes: radiance, denoised, moments_curr, moments_prev, gbuffer, gbuffer_prev, motion,
        history_prev, history_curr (tensors or None) and history_valid."""
        f = make_frame(self.storage, self.moment_format, **planes)
        check(lib().svgf_denoise(self.handle, ctypes.byref(f), _stream_handle(stream)), "svgf_denoise")

    # -- resident streams / batched form --
    def stream_open(self, w, h) -> int:
        sid = ctypes.c_uint32()
        check(lib().svgf_stream_open(self.handle, w, h, ctypes.byref(sid)), "svgf_stream_open")
        return sid.value

    def stream_close(self, sid):
        check(lib().svgf_stream_close(self.handle, sid), "svgf_stream_close")

    def stream_reset(self, sid, stream=None):
        check(lib().svgf_stream_reset(self.handle, sid, _stream_handle(stream)), "svgf_stream_reset")

    def stream_state(self, sid) -> StreamState:
        s = StreamState()
        check(lib().svgf_stream_state_get(self.handle, sid, ctypes.byref(s)), "svgf_stream_state_get")
        return s

    def budget_used(self) -> int:
        b = ctypes.c_uint64()
        check(lib().svgf_budget_used(self.handle, ctypes.byref(b)), "svgf_budget_used")
        return b.value

    def stage_times(self):
        """GPU milliseconds of each stage of the last call (needs PROFILE_STAGES)."""
        buf = (ctypes.c_float * 16)()
        n = ctypes.c_uint32()
        check(lib().svgf_stage_times(self.handle, buf, 16, ctypes.byref(n)), "svgf_stage_times")
        return [buf[i] for i in range(n.value)]

    def denoise_batch(self, items, stream=None):
        """items: list of dicts {stream_id, radiance, denoised, gbuffer?, motion?} (tensors) and
        optional request flags (REQUEST_KEEP_GBUFFER)."""
        st = self.storage
        arr = (Request * max(1, len(items)))()
        for i, it in enumerate(items):
            r = arr[i]
            r.stream_id = it["stream_id"]
            r.flags = it.get("flags", 0)
            r.radiance = plane_of(it["radiance"], fmt4(st))
            r.denoised = plane_of(it["denoised"], fmt4(st))
            r.gbuffer = plane_of(it.get("gbuffer"), FMT_RGBA32F)
            r.motion = plane_of(it.get("motion"), FMT_RG32F)
        check(lib().svgf_denoise_batch(self.handle, arr, len(items), _stream_handle(stream)), "svgf_denoise_batch")

    def batch_last(self) -> int:
        b = ctypes.c_uint64()
        check(lib().svgf_batch_last(self.handle, ctypes.byref(b)), "svgf_batch_last")
        return b.value

    def batch_status(self, batch: int, wait: bool = True) -> int:
        """svgf_status of a batch (0 ok, 9 timeout, 10 not ready)."""
        return lib().svgf_batch_status(self.handle, batch, int(wait))


class Server:
    """Continuous-batching server over a Context's resident streams (svgf_server_*)."""

    def __init__(self, ctx: Context, lanes=2, max_batch_frames=None, max_batch_pixels=None, max_wait_us=0):
        d = ServerDesc(lanes, max_batch_frames or ctx.desc.max_batch_frames,
                       max_batch_pixels or ctx.desc.max_batch_pixels, max_wait_us, 0)
        h = ctypes.c_void_p()
        check(lib().svgf_server_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)), "svgf_server_create")
        self.ctx, self.handle = ctx, h

    def close(self):
        if self.handle:
            check(lib().svgf_server_destroy(self.handle), "svgf_server_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stream_open(self, w, h) -> int:
        sid = ctypes.c_uint32()
        check(lib().svgf_server_stream_open(self.handle, w, h, ctypes.byref(sid)), "svgf_server_stream_open")
        return sid.value

    def stream_close(self, sid):
        check(lib().svgf_server_stream_close(self.handle, sid), "svgf_server_stream_close")

    def submit(self, item: dict, tag: int):
        st = self.ctx.storage
        r = Request()
        r.stream_id = item["stream_id"]
        r.flags = item.get("flags", 0)
        r.radiance = plane_of(item["radiance"], fmt4(st))
        r.denoised = plane_of(item["denoised"], fmt4(st))
        r.gbuffer = plane_of(item.get("gbuffer"), FMT_RGBA32F)
        r.motion = plane_of(item.get("motion"), FMT_RG32F)
        check(lib().svgf_server_submit(self.handle, ctypes.byref(r), tag), "svgf_server_submit")

    def poll(self, wait_us=0, capacity=256):
        buf = (Completion * capacity)()
        n = ctypes.c_uint32()
        check(lib().svgf_server_poll(self.handle, buf, capacity, ctypes.byref(n), wait_us), "svgf_server_poll")
        return [buf[i] for i in range(n.value)]

    def drain(self):
        check(lib().svgf_server_drain(self.handle), "svgf_server_drain")

    def simulate(self, *, streams, classes, class_inputs, rate_hz, duration_s, churn_per_s, seed=0):
        """classes: list of (w, h); class_inputs: list (per class) of lists of input dicts
        (radiance, gbuffer, motion tensors)."""
        st = self.ctx.storage
        ipc = len(class_inputs[0])
        arr = (Request * (len(classes) * ipc))()
        for c, ins in enumerate(class_inputs):
            assert len(ins) == ipc
            for k, it in enumerate(ins):
                r = arr[c * ipc + k]
                r.flags = it.get("flags", 0)
                r.radiance = plane_of(it["radiance"], fmt4(st))
                r.gbuffer = plane_of(it.get("gbuffer"), FMT_RGBA32F)
                r.motion = plane_of(it.get("motion"), FMT_RG32F)
        cw = (ctypes.c_uint32 * len(classes))(*[w for w, _ in classes])
        ch = (ctypes.c_uint32 * len(classes))(*[h for _, h in classes])
        d = SimDesc(streams, len(classes), cw, ch, arr, ipc, 0, rate_hz, duration_s, churn_per_s, seed)
        res = SimResult()
        check(lib().svgf_server_simulate(self.handle, ctypes.byref(d), ctypes.byref(res)), "svgf_server_simulate")
        return res.as_dict()


def _stream_handle(stream):
    if stream is None:
        torch = _torch()
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def device_tensor_from_plane(plane: Plane, h: int, w: int, channels: int, dtype):
    """Copy a library-owned plane (e.g. resident stream state) into a new torch tensor."""
    torch = _torch()
    out = torch.empty((h, w, channels), dtype=dtype, device="cuda")
    elem = out.element_size() * channels
    import ctypes as _c
    hip = _hip()
    rc = hip.hipMemcpy2D(_c.c_void_p(out.data_ptr()), _c.c_size_t(w * elem), _c.c_void_p(plane.ptr),
                         _c.c_size_t(plane.pitch_bytes), _c.c_size_t(w * elem), _c.c_size_t(h), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy2D failed: {rc}")
    return out


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = ctypes.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy2D.restype = ctypes.c_int
    return _hip_lib


# ---------------------------------------------------------------------------------------------
# agent frames on the wire (RGB888 + normalised depth) and the readback ring
# ---------------------------------------------------------------------------------------------
def wire_default_params() -> WireParams:
    p = WireParams()
    check(lib().svgf_wire_default_params(ctypes.byref(p)), "svgf_wire_default_params")
    return p


class Wire:
    """Batched wire encoder: denoised colour + G-buffer depth -> RGB888 + depth in [0, 1]."""

    def __init__(self, max_frames=256, device=0):
        self.handle = ctypes.c_void_p()
        check(lib().svgf_wire_create(device, max_frames, ctypes.byref(self.handle)), "svgf_wire_create")

    def close(self):
        if self.handle:
            check(lib().svgf_wire_destroy(self.handle), "svgf_wire_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode(self, items, params=None, stream=None):
        """items: dicts with colour (h, w, 4) fp16/fp32, gbuffer (h, w, 4) fp32, rgb (uint8 tensor of
        3 w h bytes), depth (float32 tensor of w h values), all on the device."""
        torch = _torch()
        arr = (WireFrame * len(items))()
        for k, it in enumerate(items):
            h, w = it["colour"].shape[:2]
            f = arr[k]
            f.width, f.height = w, h
            f.colour = plane_of(it["colour"], FMT_RGBA16F if it["colour"].dtype == torch.float16 else FMT_RGBA32F)
            f.gbuffer = plane_of(it["gbuffer"], FMT_RGBA32F)
            f.rgb = it["rgb"].data_ptr()
            f.depth = it["depth"].data_ptr()
        p = params or wire_default_params()
        check(lib().svgf_wire_encode(self.handle, arr, len(items), ctypes.byref(p), _stream_handle(stream)),
              "svgf_wire_encode")


class Readback:
    """Pinned host ring of `slots` buffers (svgf_readback_*)."""

    def __init__(self, slot_bytes, slots=3, device=0):
        self.handle = ctypes.c_void_p()
        check(lib().svgf_readback_create(device, slot_bytes, slots, ctypes.byref(self.handle)), "svgf_readback_create")

    def close(self):
        if self.handle:
            check(lib().svgf_readback_destroy(self.handle), "svgf_readback_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push(self, device_ptr, nbytes, stream=None) -> int:
        seq = ctypes.c_uint64()
        check(lib().svgf_readback_push(self.handle, ctypes.c_void_p(device_ptr), nbytes, _stream_handle(stream),
                                       ctypes.byref(seq)), "svgf_readback_push")
        return seq.value

    def view(self, seq, wait=True):
        """numpy uint8 view of the host copy of push `seq` (valid until the slot is reused)."""
        ptr, n = ctypes.c_void_p(), ctypes.c_uint64()
        check(lib().svgf_readback_view(self.handle, seq, 1 if wait else 0, ctypes.byref(ptr), ctypes.byref(n)),
              "svgf_readback_view")
        return np.ctypeslib.as_array((ctypes.c_uint8 * n.value).from_address(ptr.value))


# ---------------------------------------------------------------------------------------------
# consumer composite (one-frame-late history + reprojected reflections)
# ---------------------------------------------------------------------------------------------
KIND_NONE, KIND_READY, KIND_REPROJECT = 0.0, 1.0, 2.0


def composite_default_params() -> CompositeParams:
    p = CompositeParams()
    check(lib().svgf_composite_default_params(ctypes.byref(p)), "svgf_composite_default_params")
    return p


class Composite:
    """Batched consumer composite (svgf_composite_*): out = (history + emissive) + reflected * F_view."""

    def __init__(self, max_frames=256, device=0):
        self.handle = ctypes.c_void_p()
        check(lib().svgf_composite_create(device, max_frames, ctypes.byref(self.handle)), "svgf_composite_create")

    def close(self):
        if self.handle:
            check(lib().svgf_composite_destroy(self.handle), "svgf_composite_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def frames(items):
        """ctypes array of svgf_composite_frame from dicts holding every plane of COMPOSITE_PLANES as a
        (h, w, 4) fp16/fp32 device tensor (reflection fp32) and prev_view_projection (16 floats,
        row-major)."""
        torch = _torch()
        arr = (CompositeFrame * len(items))()
        for k, it in enumerate(items):
            h, w = it["history"].shape[:2]
            f = arr[k]
            f.width, f.height = w, h
            f.prev_view_projection[:] = [float(v) for v in np.asarray(it["prev_view_projection"], np.float32).reshape(16)]
            for name in COMPOSITE_PLANES:
                t = it[name]
                setattr(f, name, plane_of(t, FMT_RGBA16F if t.dtype == torch.float16 else FMT_RGBA32F))
        return arr

    def run(self, items, params=None, stream=None, frames=None):
        arr = frames if frames is not None else self.frames(items)
        p = params or composite_default_params()
        check(lib().svgf_composite_run(self.handle, arr, len(arr), ctypes.byref(p), _stream_handle(stream)),
              "svgf_composite_run")


# ---------------------------------------------------------------------------------------------
# producer pack (lighting terms + position + octahedral normal -> radiance, G-buffer, motion)
# ---------------------------------------------------------------------------------------------
NO_MOTION = -1.0e6


class Pack:
    """Batched producer pack (svgf_pack_*)."""

    def __init__(self, max_frames=256, device=0):
        self.handle = ctypes.c_void_p()
        check(lib().svgf_pack_create(device, max_frames, ctypes.byref(self.handle)), "svgf_pack_create")

    def close(self):
        if self.handle:
            check(lib().svgf_pack_destroy(self.handle), "svgf_pack_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def frames(items):
        """ctypes array of svgf_pack_frame from dicts holding view_projection and
        prev_view_projection (16 floats, row-major), position (h, w, 4) fp32, normal_oct (h, w)
        int32/uint32, indirect / direct_diffuse / direct_specular (h, w, 4) fp16 or fp32, and the
        outputs radiance (h, w, 4) fp16/fp32, gbuffer (h, w, 4) fp32, motion (h, w, 2) fp32."""
        torch = _torch()
        arr = (PackFrame * len(items))()
        for k, it in enumerate(items):
            h, w = it["position"].shape[:2]
            f = arr[k]
            f.width, f.height = w, h
            f.view_projection[:] = [float(v) for v in np.asarray(it["view_projection"], np.float32).reshape(16)]
            f.prev_view_projection[:] = [float(v) for v in np.asarray(it["prev_view_projection"], np.float32).reshape(16)]
            f.position = plane_of(it["position"], FMT_RGBA32F)
            f.normal_oct = plane_of(it["normal_oct"], FMT_R32UI)
            for name in ("indirect", "direct_diffuse", "direct_specular", "radiance"):
                t = it[name]
                setattr(f, name, plane_of(t, FMT_RGBA16F if t.dtype == torch.float16 else FMT_RGBA32F))
            f.gbuffer = plane_of(it["gbuffer"], FMT_RGBA32F)
            f.motion = plane_of(it["motion"], FMT_RG32F)
        return arr

    def run(self, items, stream=None, frames=None):
        arr = frames if frames is not None else self.frames(items)
        check(lib().svgf_pack_run(self.handle, arr, len(arr), _stream_handle(stream)), "svgf_pack_run")


# ---------------------------------------------------------------------------------------------
# frame profiler (the reference's per-pass EMA + periodic log, on hipEvents)
# ---------------------------------------------------------------------------------------------
class Profiler:
    """svgf_profiler_*: per-pass start / end events per frame slot, EMA of the durations, a log
    line every `log_interval` frames. device=-1: host samples only (sample())."""

    def __init__(self, names=("svgf",), slots=2, ema_alpha=0.05, log_interval=60, first_frame=16, device=0):
        d = ProfilerDesc()
        check(lib().svgf_profiler_default_desc(ctypes.byref(d)), "svgf_profiler_default_desc")
        self._names = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        d.passes, d.slots, d.ema_alpha, d.log_interval = len(names), slots, ema_alpha, log_interval
        d.first_frame, d.device = first_frame, device
        d.names = ctypes.cast(self._names, ctypes.POINTER(ctypes.c_char_p))
        self.passes = len(names)
        self.handle = ctypes.c_void_p()
        check(lib().svgf_profiler_create(ctypes.byref(d), ctypes.byref(self.handle)), "svgf_profiler_create")

    def close(self):
        if self.handle:
            check(lib().svgf_profiler_destroy(self.handle), "svgf_profiler_destroy")
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin_frame(self) -> bool:
        logged = ctypes.c_int32()
        check(lib().svgf_profiler_begin_frame(self.handle, ctypes.byref(logged)), "svgf_profiler_begin_frame")
        return bool(logged.value)

    def pass_begin(self, k, stream=None):
        check(lib().svgf_profiler_pass_begin(self.handle, k, _stream_handle(stream)), "svgf_profiler_pass_begin")

    def pass_end(self, k, stream=None):
        check(lib().svgf_profiler_pass_end(self.handle, k, _stream_handle(stream)), "svgf_profiler_pass_end")

    def sample(self, k, ns):
        check(lib().svgf_profiler_sample(self.handle, k, float(ns)), "svgf_profiler_sample")

    def end_frame(self):
        check(lib().svgf_profiler_end_frame(self.handle), "svgf_profiler_end_frame")

    def ema(self):
        """Per-pass averages in ns (numpy float32), or None before the first readback."""
        out = np.zeros(self.passes, np.float32)
        n = ctypes.c_uint32()
        st = lib().svgf_profiler_ema(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.passes,
                                     ctypes.byref(n))
        if st == 10:
            return None
        check(st, "svgf_profiler_ema")
        return out

    def log_line(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        check(lib().svgf_profiler_log_line(self.handle, buf, len(buf)), "svgf_profiler_log_line")
        return buf.value.decode()
