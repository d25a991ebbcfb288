"""Hazard assumptions of the inline-assembly DPP derivative blocks.

The element derivatives (ca_device.h, ca_dquad / ca_dror) are inline assembly
blocks of v_mul + 3 v_fmac_f32_dpp preceded by one s_nop 1.  The compiler's
hazard recognizer does not look inside inline assembly, so the blocks carry
their own wait states, which cover a VALU write of the moved VGPR (2 wait
states).  They do not cover the other DPP hazard, a VALU write of EXEC ahead
of a DPP instruction (5 wait states): this test compiles the library for
gfx950 and checks that no kernel writes EXEC from a VALU instruction, and that
each block starts with its wait.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "climaatmos.jl_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def device_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("asm") / "unity.s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--offload-device-only", "-S",
           os.path.join(CSRC, "ca_unity.hip"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = open(out).read().split("\n")
    return [l.strip() for l in lines if l.startswith("\t") and l.strip() and not l.strip().startswith((".", ";"))]


def test_no_valu_writes_exec(device_asm):
    bad = [l for l in device_asm
           if l.startswith("v_") and (l.split()[0].startswith("v_cmpx") or re.match(r"v_\S+\s+exec\b", l))]
    assert not bad, bad[:5]


def test_each_dpp_block_starts_with_its_wait(device_asm):
    n_blocks = 0
    for i, l in enumerate(device_asm):
        if l.startswith("v_fmac_f32_dpp"):
            # the block is [s_nop 1, v_mul, fmac, fmac, fmac]: walk back to the multiply
            j = i - 1
            while device_asm[j].startswith("v_fmac_f32_dpp"):
                j -= 1
            assert device_asm[j].startswith(("v_mul_f32_dpp", "v_mul_f32_e32")), device_asm[j - 1:i + 1]
            assert device_asm[j - 1] == "s_nop 1", device_asm[j - 1:i + 1]
            if device_asm[i - 1].startswith("v_mul"):
                n_blocks += 1
    assert n_blocks > 0
