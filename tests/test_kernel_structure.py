"""Structure of the generated gfx950 code (CPU only: hipcc -S): the persistent batched kernel's
barriers, register spills, and the absence of packed FP32 instructions.

Regression test for a hang: written as two `tid == 0` blocks on either side of the ticket
loop's back edge (publish a tile; take the next ticket), the loop was restructured by the
compiler into two nested loops whose inner loop is left by thread 0 alone. The other lanes then
passed the loop's first barrier without it and reran the previous ticket forever. Every
workgroup barrier of the ticket loop must therefore sit in the loop itself (depth 1), and no loop
that contains a barrier may have an exec-mask latch (a divergent exit).
"""
from __future__ import annotations

import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not pathlib.Path(HIPCC).exists() and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "svgf_kernels.s"
    # the library's own device flags (tron-grid_amd/Makefile)
    subprocess.run([HIPCC if pathlib.Path(HIPCC).exists() else "hipcc", "--offload-arch=gfx950", "-O3", "-std=c++20",
                    "-fPIC", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
                    "-o", str(out), str(ROOT / "tron-grid_amd" / "csrc" / "svgf_kernels.hip")],
                   check=True, capture_output=True)
    return out.read_text()


def _functions(text, prefix):
    for m in re.finditer(r"^(" + prefix + r"\S*):", text, re.M):
        i = m.start()
        yield m.group(1), text[i:text.index(".Lfunc_end", i)].splitlines()


def _loop_of_blocks(lines):
    """Per line: (loop header, depth) of the basic block it belongs to, from the assembler's
    loop annotations (None outside loops)."""
    cur, out = None, []
    for n, line in enumerate(lines):
        t = line.strip()
        if t.startswith(".LBB") or t.startswith("; %bb."):
            name = t.split(":")[0].lstrip(".").replace("; %", "").lstrip("L")
            nxt = lines[n + 1].strip() if n + 1 < len(lines) else ""
            hdr = re.search(r"=>\s*This (Inner )?Loop Header: Depth=(\d+)", t) or \
                (re.search(r"=>\s*This (Inner )?Loop Header: Depth=(\d+)", nxt) if nxt.startswith(";") else None)
            inl = re.search(r"in Loop: Header=(BB\d+_\d+) Depth=(\d+)", t)
            cur = (name, int(hdr.group(2))) if hdr else ((inl.group(1), int(inl.group(2))) if inl else None)
        out.append(cur)
    return out


def test_persistent_ticket_loop_barriers_are_not_divergent(asm):
    kernels = list(_functions(asm, r"_ZN4svgf12_GLOBAL__N_19k_persist"))
    # REFERENCE fp32/fp16, CANONICAL fp32/fp16 x raw/compensated moments, fp16 wide state
    assert len(kernels) == 7
    for name, lines in kernels:
        loops = _loop_of_blocks(lines)
        latches = {loops[n][0] for n, l in enumerate(lines) if "s_andn2_b64 exec, exec" in l and loops[n]}
        barriers = [loops[n] for n, l in enumerate(lines) if "s_barrier" in l]
        in_loop = [b for b in barriers if b is not None]
        assert len(in_loop) >= 3, name  # ticket, dependency wait, publish (+ the stages' own)
        assert len({b[0] for b in in_loop}) == 1, (name, in_loop)  # all in the one ticket loop
        for b in in_loop:
            assert b[1] == 1, (name, b)
            assert b[0] not in latches, (name, b)


def test_hot_kernels_do_not_spill(asm):
    """No scratch (register spills) in any kernel, the persistent ones included."""
    meta = asm[asm.index("amdhsa.kernels"):]
    for blk in meta.split("  - .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk).group(1)
        scratch = int(re.search(r"\.private_segment_fixed_size:\s+(\d+)", blk).group(1))
        assert scratch == 0, name


# the library's device flags (tron-grid_amd/Makefile)
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "--cuda-device-only", "-S"]
PACKED_F32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def _device_asm(src, out, flags):
    hipcc = HIPCC if pathlib.Path(HIPCC).exists() else "hipcc"
    lang = ["-x", "hip"] if src.suffix == ".cpp" else []
    subprocess.run([hipcc] + flags + lang + ["-o", str(out), str(src)], check=True, capture_output=True)
    return out.read_text()


def test_no_packed_fp32_instructions(asm, tmp_path):
    """Every HIP source of the library, built with the library's flags, contains no
    v_pk_add_f32 / v_pk_mul_f32 / v_pk_fma_f32 (the -packed-fp32-ops target feature applies to
    the device pass; the warning hipcc prints about it comes from the host pass). Control: the
    kernels built without the feature do contain them, so the check can fail."""
    if not pathlib.Path(HIPCC).exists() and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    assert not PACKED_F32.search(asm)
    csrc = ROOT / "tron-grid_amd" / "csrc"
    for src in sorted(csrc.glob("*.hip")) + sorted(csrc.glob("*.cpp")):
        if src.name == "svgf_kernels.hip":
            continue
        text = _device_asm(src, tmp_path / (src.name + ".s"), DEVICE_FLAGS)
        assert not PACKED_F32.search(text), src.name
    flags = [f for f in DEVICE_FLAGS if f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")]
    control = _device_asm(csrc / "svgf_kernels.hip", tmp_path / "control.s", flags)
    assert PACKED_F32.search(control)
