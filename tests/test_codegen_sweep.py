"""Code-generation checks of the sparse all-pairs kernel (CPU only).

k_gain_sparse accumulates a row of X in dynamic LDS and, when every G2 list
weight is in [0, 16383], uses a list word shifted right by 14 as the LDS byte
address of its target.  That is correct only while the row accumulator starts
at LDS address 0, i.e. while the kernel declares no static LDS.  Its speed
depends on two 1024-thread workgroups fitting on a CU; on gfx950 that held at
72 scalar registers and failed at 80 (half the resident waves, 1.84 ms against
1.40 ms for one sweep at n = 18,524), although the compiler's own occupancy
estimate was the same for both.  This test compiles bm_swap.hip to gfx950
assembly and checks both properties for every instantiation.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def kernel_descriptors(asm_text, name):
    """{mangled kernel: {field: int}} for the .amdhsa_kernel blocks whose name contains `name`."""
    out = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm_text, re.S):
        if name in m.group(1):
            out[m.group(1)] = {k: int(v) for k, v in re.findall(r"\.amdhsa_(\w+) (\d+)", m.group(2))}
    return out


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_sparse_gain_kernel_layout_and_registers(tmp_path):
    out = tmp_path / "bm_swap.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
           os.path.join(ROOT, "brain-match_amd", "csrc", "bm_swap.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    desc = kernel_descriptors(out.read_text(), "k_gain_sparse")
    assert len(desc) == 8  # <hubs, neg, small> for every combination
    for k, d in desc.items():
        assert d["group_segment_fixed_size"] == 0, k
        assert d["next_free_sgpr"] <= 72, (k, d["next_free_sgpr"])
        assert d["next_free_vgpr"] <= 64, (k, d["next_free_vgpr"])
