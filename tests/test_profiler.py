"""The frame profiler (CPU only): the oracle of the reference's profiler loop
(oracle/profiler_oracle.c, src/main.cpp:2910-2951, 3043, 3096, 3148) pinned by what the loop
fixes, and the library's profiler (host-sample mode, the same bookkeeping the hipEvent mode
runs) equal to the oracle bit for bit.

Pins: the first readback happens `slots` iterations after the first frame and copies that
frame's samples exactly; a step in the samples reaches the average exactly `slots` iterations
later and then decays geometrically by (1 - alpha) per frame (float64 closed form); log lines
come exactly at the iterations whose frame counter (first + i) is a multiple of the interval,
from the first readback on; the line formats each average in ms with C's %f.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import svgf_amd as S


def test_first_readback_copies_the_first_frame():
    rng = np.random.default_rng(1)
    s = rng.uniform(1e4, 1e6, (10, 3)).astype(np.float32)
    for slots in (1, 2, 3):
        ema, logged = O.profiler_run(s, slots=slots)
        assert np.all(ema[:slots] == 0.0)
        assert np.array_equal(ema[slots], s[0])


def test_step_reaches_the_average_after_the_slot_delay_and_decays_geometrically():
    a, b, k, slots, alpha = 1.0e5, 3.0e5, 20, 2, 0.05
    s = np.full((200, 1), a, np.float32)
    s[k:] = b
    ema, _ = O.profiler_run(s, slots=slots, alpha=alpha)
    # before the step's sample is read back, the average stays near a
    assert np.all(np.abs(ema[slots:k + slots, 0] - a) <= a * 1e-5)
    n = np.arange(k + slots, 200)
    want = b + (a - b) * (1.0 - alpha) ** (n - (k + slots) + 1)
    assert np.all(np.abs(ema[n, 0] - want) <= 1e-4 * b)


def test_log_cadence():
    s = np.ones((400, 2), np.float32) * 5e5
    for first, interval, slots in [(16, 60, 2), (0, 60, 2), (16, 7, 3), (59, 60, 1)]:
        _, logged = O.profiler_run(s, slots=slots, interval=interval, first=first)
        want = np.array([(i >= slots) and ((first + i) % interval == 0) for i in range(400)])
        assert np.array_equal(logged, want), (first, interval, slots)
    _, logged = O.profiler_run(s)
    assert np.flatnonzero(logged)[:3].tolist() == [44, 104, 164]  # frame counter 60, 120, 180


def _library_run(samples, slots=2, alpha=0.05, interval=60, first=16, names=None):
    frames, passes = samples.shape
    names = names or [f"p{k}" for k in range(passes)]
    emas, logged, lines = [], [], []
    with S.Profiler(names, slots=slots, ema_alpha=alpha, log_interval=interval, first_frame=first, device=-1) as p:
        for i in range(frames):
            logged.append(p.begin_frame())
            for k in range(passes):
                p.sample(k, samples[i, k])
            p.end_frame()
            e = p.ema()
            emas.append(np.zeros(passes, np.float32) if e is None else e)
            lines.append(p.log_line())
    return np.array(emas), np.array(logged), lines


@pytest.mark.parametrize("slots,alpha,interval,first", [(2, 0.05, 60, 16), (1, 0.2, 7, 0), (3, 0.05, 10, 5)])
def test_library_profiler_equals_the_oracle(slots, alpha, interval, first):
    rng = np.random.default_rng(slots)
    s = rng.lognormal(12.0, 0.5, (300, 4)).astype(np.float32)
    ema_o, log_o = O.profiler_run(s, slots=slots, alpha=alpha, interval=interval, first=first)
    ema_l, log_l, lines = _library_run(s, slots, alpha, interval, first)
    assert np.array_equal(ema_l.view(np.uint32), ema_o.view(np.uint32))
    assert np.array_equal(log_l, log_o)
    i = int(np.flatnonzero(log_o)[0])
    want = "GPU times (ms):" + "".join(f" p{k}=%f" % float(np.float32(ema_o[i, k]) / np.float32(1e6))
                                       for k in range(4))
    assert lines[i] == want


def test_log_line_format_matches_the_reference_names():
    names = ["frame", "mesh", "svgf", "vol_inject", "vol_filter", "vol_composite", "bloom_down", "bloom_up", "post"]
    s = np.tile(np.arange(1, 10, dtype=np.float32) * 1e5, (50, 1))
    _, logged, lines = _library_run(s, names=names)
    i = int(np.flatnonzero(logged)[0])
    assert lines[i] == ("GPU times (ms): frame=0.100000 mesh=0.200000 svgf=0.300000 vol_inject=0.400000 "
                        "vol_filter=0.500000 vol_composite=0.600000 bloom_down=0.700000 bloom_up=0.800000 "
                        "post=0.900000")


def test_profiler_argument_checks():
    L = S.lib()
    import ctypes

    d = S.ProfilerDesc()
    assert L.svgf_profiler_default_desc(ctypes.byref(d)) == 0
    assert (d.passes, d.slots, d.log_interval, d.first_frame) == (1, 2, 60, 16) and abs(d.ema_alpha - 0.05) < 1e-9
    h = ctypes.c_void_p()
    for field, bad in (("passes", 0), ("passes", 33), ("slots", 0), ("slots", 9), ("log_interval", 0),
                       ("ema_alpha", 1.5), ("device", -2)):
        e = S.ProfilerDesc()
        L.svgf_profiler_default_desc(ctypes.byref(e))
        e.device = -1
        setattr(e, field, bad)
        assert L.svgf_profiler_create(ctypes.byref(e), ctypes.byref(h)) == 1, field
    with S.Profiler(["a"], device=-1) as p:
        with pytest.raises(S.SvgfError):
            p.sample(0, 1.0)  # outside a frame
        p.begin_frame()
        with pytest.raises(S.SvgfError):
            p.begin_frame()  # frames do not nest
        with pytest.raises(S.SvgfError):
            p.pass_begin(0, stream=0)  # no events in host mode
        with pytest.raises(S.SvgfError):
            p.sample(1, 1.0)  # no such pass
        p.end_frame()
        assert p.ema() is None and p.log_line() == ""
