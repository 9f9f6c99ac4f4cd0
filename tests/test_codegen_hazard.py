"""Code-generation check of the single-thread state kernels (CPU only).

A uniform load of a device word may be compiled to a scalar-cache load.  If
the same thread then overwrites that word with a vector store whose value does
not depend on the load, the store can land before the load returns, and the
load sees the new value.  The refinement state kernels (bm_refine.hip) read
and rewrite one state struct; they read it through a volatile pointer so that
the reads are vector loads.  This test compiles the file to gfx950 assembly and
scans every kernel for a vector store that overlaps a still-pending scalar
load from the same base register pair (no ``s_waitcnt lgkmcnt(0)`` between).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

_SLOAD = re.compile(r"\s*s_load_dword(x\d+)?\s+(s\[\d+:\d+\]|s\d+),\s*(s\[\d+:\d+\]),\s*(0x[0-9a-f]+|\d+)")
_STORE = re.compile(r"\s*global_store_(dword|dwordx2|dwordx3|dwordx4|short|byte)\s+v\d+,\s*(v\[\d+:\d+\]|v\d+),"
                    r"\s*(s\[\d+:\d+\])(?:\s+offset:(-?\d+))?")
_WAIT = re.compile(r"\s*s_waitcnt\s+(.*)")
_LOAD_BYTES = {None: 4, "x2": 8, "x3": 12, "x4": 16, "x8": 32, "x16": 64}
_STORE_BYTES = {"dword": 4, "dwordx2": 8, "dwordx3": 12, "dwordx4": 16, "short": 2, "byte": 1}


def pending_load_overwrites(asm_lines):
    """(function, line) of each store over a pending scalar load, in a linear scan."""
    fn, pending, hits = None, [], []
    for ln, line in enumerate(asm_lines, 1):
        if re.match(r"^[A-Za-z_][\w.]*:", line):
            fn, pending = line.split(":")[0], []
            continue
        m = _SLOAD.match(line)
        if m:
            base, off, dst = m.group(3), int(m.group(4), 0), m.group(2)
            pending = [p for p in pending if p[0] != dst]  # the base register is redefined
            pending.append((base, off, off + _LOAD_BYTES[m.group(1)]))
            continue
        m = _WAIT.match(line)
        if m and ("lgkmcnt(0)" in m.group(1) or m.group(1).strip() == "0"):
            pending = []
            continue
        m = _STORE.match(line)
        if m:
            base, lo = m.group(3), int(m.group(4) or 0)
            hi = lo + _STORE_BYTES[m.group(1)]
            if any(b == base and a0 < hi and lo < a1 for b, a0, a1 in pending):
                hits.append((fn, ln))
    return hits


def test_scanner_finds_the_pattern():
    asm = """k:
\ts_load_dwordx2 s[2:3], s[0:1], 0x0
\ts_waitcnt lgkmcnt(0)
\ts_load_dwordx2 s[4:5], s[2:3], 0x50
\tglobal_store_dwordx4 v4, v[0:3], s[2:3] offset:80
\ts_waitcnt lgkmcnt(0)
\tglobal_store_dwordx2 v4, v[0:1], s[2:3] offset:80
""".splitlines()
    assert pending_load_overwrites(asm) == [("k", 5)]


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_refine_state_kernels_keep_load_store_order(tmp_path):
    out = tmp_path / "bm_refine.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
           os.path.join(ROOT, "brain-match_amd", "csrc", "bm_refine.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    lines = out.read_text().splitlines()
    assert any("k_state_init" in ln for ln in lines)
    assert pending_load_overwrites(lines) == []
