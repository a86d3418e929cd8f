"""The golden fixtures of tests/golden against the C oracle (CPU only).

The fixtures come from tests/golden/svgf_restate.py, an FP64 restatement of both modes written
apart from the C oracle (whole-image numpy arrays with clamped index grids, not per-pixel
loops). The oracle must reproduce them to rounding. The fixtures must also see each plausible
change of a CANONICAL rule: the two documented alternatives (Schied's +-1 variance prefilter,
epsilon outside the square root) and perturbed constants, all planted in the oracle.
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

import oracle as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_fixtures  # noqa: E402


def _rel(a, b, floor=1e-3):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b) / np.maximum(np.abs(b), floor)))


@pytest.fixture(scope="module")
def ref():
    return np.load(GOLDEN / "reference.npz")


@pytest.fixture(scope="module")
def can():
    return np.load(GOLDEN / "canonical.npz")


def test_fixtures_are_what_the_script_writes(ref, can):
    """The committed files are the script's output (inputs bit for bit; outputs to rounding,
    which may differ between numpy builds in the last bits of exp and pow)."""
    for want, got in ((ref, make_fixtures.reference_fixture()), (can, make_fixtures.canonical_fixture())):
        assert sorted(want.files) == sorted(got)
        for k in want.files:
            if want[k].dtype == np.float64:
                np.testing.assert_allclose(got[k], want[k], rtol=1e-12, atol=1e-15, err_msg=k)
            else:
                assert np.array_equal(got[k], want[k]), k


def test_reference_oracle_matches_fixtures(ref, oracle_lib):
    passes = int(ref["passes"])
    for i, (w, h) in enumerate(ref["shapes"]):
        mom = mom32 = None
        for f in range(ref[f"s{i}_radiance"].shape[0]):
            rad = ref[f"s{i}_radiance"][f]
            mc, den = O.r_frame_f64(rad.astype(np.float64), mom, O.reference_params(passes),
                                    history_valid=mom is not None)
            assert _rel(mc, ref[f"s{i}_moments"][f]) <= 1e-13, (i, f)
            assert _rel(den, ref[f"s{i}_denoised"][f]) <= 1e-13, (i, f)
            # the float restatement of the shader stays within its 1e-5 parity tolerance
            m32, d32, _, _ = O.r_frame_f32(rad, mom32, O.reference_params(passes), history_valid=mom32 is not None)
            assert _rel(d32, ref[f"s{i}_denoised"][f]) <= 1e-5, (i, f)
            mom, mom32 = mc, m32


def _canonical_errors(can, form, params_of):
    """Worst relative error of the oracle's history colour, moments and output against the
    fixtures of a moment form, each stream replayed by the oracle from its own state."""
    worst = 0.0
    lengths_equal = True
    for i, (w, h) in enumerate(can["shapes"]):
        hist = mom = gp = None
        for f in range(can[f"s{i}_radiance"].shape[0]):
            o = O.c_frame_f64(can[f"s{i}_radiance"][f], can[f"s{i}_gbuffer"][f], gp, can[f"s{i}_motion"][f], hist,
                              mom, params_of(), history_valid=f > 0)
            worst = max(worst, _rel(o["history"][..., :3], can[f"s{i}_{form}_history"][f][..., :3]),
                        _rel(o["moments"], can[f"s{i}_{form}_moments"][f]),
                        _rel(o["denoised"], can[f"s{i}_{form}_denoised"][f]))
            lengths_equal &= np.array_equal(o["history"][..., 3], can[f"s{i}_{form}_history"][f][..., 3])
            hist, mom, gp = o["history"], o["moments"], can[f"s{i}_gbuffer"][f]
    return worst, lengths_equal


@pytest.mark.parametrize("form", ["raw", "comp"])
def test_canonical_oracle_matches_fixtures(can, oracle_lib, form):
    worst, lengths = _canonical_errors(can, form, lambda: O.canonical_params(5, compensated=form == "comp"))
    assert lengths
    assert worst <= 1e-12, worst


def _planted(change):
    def make():
        p = O.canonical_params(5, compensated=True)
        change(p)
        return p
    return make


PLANTED = {
    "prefilter_at_unit_offsets": lambda p: setattr(p, "flags", p.flags | 0x200),
    "eps_outside_the_root": lambda p: setattr(p, "flags", p.flags | 0x400),
    "alpha_moments_0.25": lambda p: setattr(p, "alpha_moments", 0.25),
    "alpha_color_0.25": lambda p: setattr(p, "alpha_color", 0.25),
    "phi_normal_64": lambda p: setattr(p, "phi_normal", 64.0),
    "phi_depth_2": lambda p: setattr(p, "phi_depth", 2.0),
    "history_cap_4": lambda p: setattr(p, "history_cap", 4),
    "depth_reject_0.005": lambda p: setattr(p, "depth_reject", 0.005),
    "normal_reject_0.9999": lambda p: setattr(p, "normal_reject", 0.9999),
    "four_iterations": lambda p: setattr(p, "iterations", 4),
}


@pytest.mark.parametrize("name", list(PLANTED))
def test_fixtures_see_a_changed_rule(can, oracle_lib, name):
    worst, lengths = _canonical_errors(can, "comp", _planted(PLANTED[name]))
    assert worst > 1e-6 or not lengths, (name, worst)


@pytest.mark.parametrize("form", ["raw", "comp"])
def test_float_port_close_to_fixtures(can, oracle_lib, form):
    """The binary32 build of the CANONICAL code (the bench's CPU baseline port, float state
    carried from frame to frame) within 5e-5 of the FP64 fixtures, history lengths exact."""
    worst = 0.0
    for i, (w, h) in enumerate(can["shapes"]):
        hist = mom = gp = None
        for f in range(can[f"s{i}_radiance"].shape[0]):
            o = O.c_frame_f32(can[f"s{i}_radiance"][f], can[f"s{i}_gbuffer"][f], gp, can[f"s{i}_motion"][f], hist, mom,
                              O.canonical_params(5, compensated=form == "comp"), history_valid=f > 0)
            want_h = can[f"s{i}_{form}_history"][f]
            assert np.array_equal(o["history"][..., 3], want_h[..., 3]), (i, f)
            worst = max(worst, _rel(o["history"][..., :3], want_h[..., :3]), _rel(o["denoised"], can[f"s{i}_{form}_denoised"][f]),
                        _rel(o["moments"], can[f"s{i}_{form}_moments"][f]))
            hist, mom, gp = o["history"], o["moments"], can[f"s{i}_gbuffer"][f]
    assert worst <= 5e-5, worst
