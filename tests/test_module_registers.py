"""Register budget of the step kernels (no GPU needed: hipRTC compiles for
gfx950 on the host, and the code object's metadata is read with the ROCm
llvm-readelf).

The select and compaction kernels run every step.  gpugo_compact_select and
gpugo_select are launched with 1024 threads, which caps a wavefront at 128
VGPRs; past that the compiler spills to scratch, and a kernel that needs
scratch made every step of small solves about 30% slower (measured on
MI355X: the 64^2-grid d = 2 solve 0.23 -> 0.32 ms).  So none of them may use
a private segment, in an unrolled, a term-parallel or a windowed module.
"""
import os
import re
import subprocess

import pytest

from gpugo_amd import problems as pb
from gpugo_amd.solver import rtc_compile_to_cache

READELF = "/opt/rocm/llvm/bin/llvm-readelf"
STEP_KERNELS = ("gpugo_compact_select", "gpugo_select", "gpugo_compact", "gpugo_root")


def kernel_meta(path):
    out = subprocess.run([READELF, "--notes", path], capture_output=True, text=True, check=True).stdout
    names = re.findall(r"\.name:\s+(\S+)", out)
    priv = [int(v) for v in re.findall(r"\.private_segment_fixed_size:\s+(\d+)", out)]
    vgpr = [int(v) for v in re.findall(r"\.vgpr_count:\s+(\d+)", out)]
    assert len(names) == len(priv) == len(vgpr) and names
    return {n: (p, v) for n, p, v in zip(names, priv, vgpr)}


@pytest.mark.skipif(not os.path.exists(READELF), reason="ROCm llvm-readelf not installed")
@pytest.mark.parametrize("kind", ["unrolled", "split", "windowed_f32"])
def test_step_kernels_use_no_scratch(kind):
    if kind == "unrolled":
        s = pb.sepcos2()
        path = rtc_compile_to_cache(pb.objective(s.family, s.d), roundings=("fast",))
    elif kind == "split":
        s = pb.chain10()
        path = rtc_compile_to_cache(pb.objective(s.family, s.d), roundings=("fast",), split=4)
    else:
        s = pb.ackley50()
        path = rtc_compile_to_cache(pb.objective_loops(s.family, s.d), roundings=("fast",), f32=True,
                                    windowed=True)
    meta = kernel_meta(path)
    for k in STEP_KERNELS:
        assert k in meta, (k, sorted(meta))
        priv, vgpr = meta[k]
        assert priv == 0, (kind, k, priv, vgpr)
        assert vgpr <= 128, (kind, k, vgpr)
