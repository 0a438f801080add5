"""Guard on the device code of the location-code helpers (CPU only: compiles to gfx950
assembly, runs nothing).

On gfx950 the optimiser's 16-bit form of locDepth's division by 3 (a 16-bit multiply
followed by byte selects) gave sporadic wrong pre-order keys -- whole wavefronts, a few
per million, varying between runs -- which scrambled node ids of large trees.  The helper
now keeps that division in 32-bit operations; this test fails if a compiler change
brings the 16-bit form back.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"

SRC = r'''
#include "common.h"
using namespace tsr;
__global__ void k_probeDepth(const u64 *locs, u32 n, u64 *keys, u32 *depths) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) { keys[i] = dfsKey(locs[i]); depths[i] = locDepth(locs[i]); }
}
'''


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_location_code_depth_has_no_16bit_multiply(tmp_path):
    src = tmp_path / "probe.hip"
    src.write_text(SRC)
    out = tmp_path / "probe.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", "-I", os.path.join(ROOT, "tsr_amd", "csrc"),
           "-o", str(out), str(src)]
    subprocess.run(cmd, check=True, capture_output=True)
    asm = out.read_text()
    body = asm[asm.index("k_probeDepth"):]
    body = body[:body.index("s_endpgm")]
    bad = re.findall(r"v_[a-z0-9_]+_(?:u16|i16|b16)(?:_e32|_e64|_sdwa)?\b", body)
    assert not bad, f"16-bit ops in the location-code helpers: {bad}"
    assert "sdwa" not in body
