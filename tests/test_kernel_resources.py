"""Register budgets of the hand-written FP64 kernels, read from the
compiler's resource-usage remarks (hipcc cross-compiles for gfx950; CPU
only).  A register array indexed by a run-time value, or a change that
pushes a kernel over its register budget, shows up as scratch memory
(spills) or as a lower occupancy long before it shows up as a slower run;
this test keeps those budgets fixed:
  * the batched product (k_apply_v), the pseudo-Hessian element columns
    (k_fcols_sf) and the element derivative products (k_dab_sf) use no
    scratch at any order or tiling;
  * the FP64 gradient (k_grad_sf) spills at most a few registers (the
    addresses of the prefetch and the point's inverse Jacobian, which it
    keeps in registers to save LDS), at most 64 bytes per lane, and the FP32
    gradient none;
  * the product for p4 with one row tile (the benchmark's kernel) keeps
    three waves per SIMD and at most 168 registers, and the gradient and
    pseudo-Hessian kernels for p2..p5 keep four.
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


def parse(remarks: str) -> dict[str, dict[str, int]]:
    out = {}
    for blk in re.split(r"remark: Function Name: ", remarks)[1:]:
        name = blk.split()[0]
        vals = {}
        for key, pat in (("vgpr", r"VGPRs: (\d+)"), ("scratch", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("occupancy", r"Occupancy \[waves/SIMD\]: (\d+)")):
            m = re.search(pat, blk)
            if m:
                vals[key] = int(m.group(1))
        out[name] = vals
    return out


@pytest.fixture(scope="module")
def usage():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    tmp = tempfile.mkdtemp(prefix="f25ru")
    try:
        cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-c",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-mllvm", "-amdgpu-mfma-vgpr-form", "-Wno-unused-result",
               "-Rpass-analysis=kernel-resource-usage",
               "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "fem25_amd", "csrc", "capi.hip"),
               "-o", os.path.join(tmp, "capi.o")]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        return parse(r.stderr)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def kernels(usage, base: str):
    # Itanium mangling: f25k::<base> is "4f25k<len><base>I"
    tag = f"4f25k{len(base)}{base}I"
    found = {n: v for n, v in usage.items() if tag in n}
    assert found, base
    return found


def nl_of(name: str) -> int:
    return int(re.search(r"ILi(\d+)E", name).group(1))


def test_no_scratch_in_product_and_element_kernels(usage):
    for base in ("k_apply_v", "k_fcols_sf", "k_dab_sf"):
        for name, v in kernels(usage, base).items():
            assert v["scratch"] == 0, (name, v)


def test_gradient_spills_bounded(usage):
    for name, v in kernels(usage, "k_grad_sf").items():
        assert v["scratch"] <= (0 if "ELi2EfE" in name else 64), (name, v)


def test_occupancy_budgets(usage):
    prod = {n: v for n, v in kernels(usage, "k_apply_v").items() if "ILi5ELi1E" in n}
    assert len(prod) == 1
    v = next(iter(prod.values()))
    assert v["vgpr"] <= 168 and v["occupancy"] >= 3, v
    for base in ("k_grad_sf", "k_fcols_sf"):
        for name, v in kernels(usage, base).items():
            if nl_of(name) <= 5:
                assert v["occupancy"] >= 4, (name, v)


def test_parser_reads_remarks():
    text = ("x.hip:1:1: remark: Function Name: _ZN4f25k9k_apply_vILi5ELi1EEEvv [-R]\n"
            "x.hip:1:1: remark:     VGPRs: 162 [-R]\nx.hip:1:1: remark:     ScratchSize [bytes/lane]: 0 [-R]\n"
            "x.hip:1:1: remark:     Occupancy [waves/SIMD]: 3 [-R]\n")
    u = parse(text)
    assert u == {"_ZN4f25k9k_apply_vILi5ELi1EEEvv": {"vgpr": 162, "scratch": 0, "occupancy": 3}}
    assert nl_of("_ZN4f25k9k_apply_vILi5ELi1EEEvv") == 5
