"""Build libnyles_hip.so in-tree with hipcc for gfx950.

    python -m nyles_amd.build          # compile (always)

The FP64 kernels reproduce the reference's operation order, so the whole
library is compiled with -ffp-contract=off; packed FP32 instructions are
disabled (-packed-fp32-ops) as required on the target machines.  Each
source is compiled to its own object in parallel (no relocatable device
code: no device function crosses a translation unit), then linked.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnyles_hip.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["nyl_api.hip", "nyl_stencil.hip", "nyl_mg.hip", "nyl_batch.hip", "nyl_coarse.hip",
           "nyl_noalloc.hip", "nyl_mgfor.hip"]
ARCH = os.environ.get("NYL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed (exit %d): %s" % (proc.returncode, cmd[-1]))


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in srcs]
    cmds = [[hipcc()] + FLAGS + inc + ["-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
