"""Kernel equivalence on the CPU (no GPU).

The tiled, fused and four-cells-per-thread kernels (and the 2.5-D stage
kernel) must give the same bits as the per-cell kernels they replace (the FP64 path is bit-identical to the
oracle through the per-cell arithmetic).  The kernel sources are compiled
for the host with g++ against a CPU stand-in of the HIP header
(tests/cpu_emu/hip/hip_runtime.h: one host thread per work-item with a real
barrier) and AddressSanitizer, and run on small grids of every periodicity,
including levels smaller than a tile and one-cell axes.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


@pytest.mark.parametrize("name", ["cell_emu", "mg_emu", "coarse_emu", "stage_emu"])
def test_kernels_match_per_cell_forms_bitwise(name, tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path / name)
    src = os.path.join(EMU, name + ".cpp")
    build = [gxx, "-std=c++17", "-O1", "-fsanitize=address", "-fno-omit-frame-pointer",
             "-I", EMU, "-I", INCLUDE, src, "-lpthread", "-o", exe]
    subprocess.run(build, check=True, timeout=900)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900)
    tail = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0 and "ALL OK" in r.stdout, tail
