"""The DPP operands of the hand-written FP64 kernels are passed through
inline assembly (v_fmac_f64_dpp / v_mov_b64_dpp with row_newbcast), which the
compiler's hazard recognizer does not look into.  gfx950 needs two wait
states between a vector-ALU write of a VGPR and a DPP read of it; this test
compiles the device code to assembly and checks that no vector-ALU
instruction writing a DPP source register comes within two wait states
before the DPP instruction (an s_nop N counts N + 1).  A label inside that
window (another path may reach the instruction) fails the check as well.
CPU only: hipcc cross-compiles for gfx950.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

REG = re.compile(r"v\[(\d+):(\d+)\]|v(\d+)\b")


def regs(op: str) -> set[int]:
    m = REG.fullmatch(op.strip())
    if not m:
        return set()
    if m.group(3) is not None:
        return {int(m.group(3))}
    return set(range(int(m.group(1)), int(m.group(2)) + 1))


def operands(line: str) -> list[str]:
    body = line.split(";")[0].strip()
    parts = body.split(None, 1)
    if len(parts) < 2:
        return []
    # drop DPP modifiers (row_newbcast:..., row_mask:..., bank_mask:...)
    ops = parts[1].split(",")
    return [o.strip().split()[0] for o in ops if o.strip()]


def hazards(asm_lines: list[str]) -> list[str]:
    found = []
    code = []   # (text, is_label)
    for ln in asm_lines:
        s = ln.strip()
        if not s or s.startswith(";") or s.startswith(".") and not s.endswith(":"):
            continue
        if s.endswith(":") and not s.startswith("\t"):
            code.append((s, True))
        elif ln.startswith("\t") and not s.startswith("."):
            code.append((s, False))
    for i, (s, lab) in enumerate(code):
        if lab or "_dpp" not in s.split()[0]:
            continue
        ops = operands(s)
        if len(ops) < 2:
            continue
        src = regs(ops[1])
        waits = 0
        j = i - 1
        while j >= 0 and waits < 2:
            t, tlab = code[j]
            if tlab:
                found.append(f"label inside the window before: {s}")
                break
            mnem = t.split()[0]
            if mnem == "s_nop":
                waits += int(t.split()[1], 0) + 1
            else:
                if mnem.startswith("v_") and not mnem.startswith("v_readlane") and not mnem.startswith("v_cmp"):
                    wops = operands(t)
                    if wops and regs(wops[0]) & src:
                        found.append(f"{t}  ->  {s}")
                waits += 1
            j -= 1
    return found


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_dpp_read_after_valu_write():
    tmp = tempfile.mkdtemp(prefix="f25dpp")
    try:
        out = os.path.join(tmp, "capi.s")
        cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-S",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-mllvm", "-amdgpu-mfma-vgpr-form", "-Wno-unused-result",
               "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "fem25_amd", "csrc", "capi.hip"),
               "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        with open(out) as f:
            lines = f.read().splitlines()
        n_dpp = sum(1 for ln in lines if "_dpp " in ln and ln.startswith("\t"))
        assert n_dpp > 1000          # the product, gradient and element kernels use it
        bad = hazards(lines)
        assert not bad, "\n".join(bad[:20])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_checker_finds_a_planted_hazard():
    """The checker itself: a VALU write of the DPP source one or two wait
    states before is reported; three wait states or an s_nop 1 clear it."""
    dpp = "\tv_fmac_f64_dpp v[86:87], v[64:65], v[78:79] row_newbcast:1 row_mask:0xf bank_mask:0xf"
    write = "\tv_mov_b64_e32 v[64:65], v[2:3]"
    other = "\tv_fmac_f64_e32 v[4:5], v[6:7], v[8:9]"
    head = ["f:", other, other]
    assert hazards(head + [write, dpp])
    assert hazards(head + [write, other, dpp])
    assert not hazards(head + [write, other, other, dpp])
    assert not hazards(head + [write, "\ts_nop 1", dpp])
    assert hazards(head + ["\tv_add_f64 v65, v1, v2", dpp])          # one half of the pair
    assert not hazards(head + ["\tv_mov_b64_e32 v[78:79], v[2:3]", other, dpp])   # src1 is not DPP
    assert hazards(head + [".LBB0_1:", dpp])                            # another path may reach it
