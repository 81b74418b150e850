"""Build libnyles_hip.so in-tree with hipcc for gfx950.

    python -m nyles_amd.build          # compile (always)

The FP64 kernels reproduce the reference's operation order, so the whole
library is compiled with -ffp-contract=off; packed FP32 instructions are
disabled (-packed-fp32-ops) as required on the target machines.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnyles_hip.so")
SOURCES = ["nyl_api.hip", "nyl_stencil.hip", "nyl_mg.hip", "nyl_batch.hip"]
ARCH = os.environ.get("NYL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed (exit %d)" % proc.returncode)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
