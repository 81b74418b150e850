"""Summaries of rocprofv3 output for the kernel tables of the design notes.

    python -m nyles_amd.profiling trace DIR [DIR ...]
        per-kernel launch counts and mean / min times from the
        `--kernel-trace --stats --output-format csv` files under DIR
    python -m nyles_amd.profiling traffic FETCH_DIR WRITE_DIR CFETCH_DIR CWRITE_DIR COPY_BYTES
        HBM bytes per launch of every kernel from a FETCH_SIZE pass and a
        WRITE_SIZE pass (`--pmc ... --output-format csv`), each scaled by what
        the same counter reads for a device copy of COPY_BYTES bytes
        (`bench.py --copy-calibration-mb`)
    python -m nyles_amd.profiling valu DIR
        VALU instructions per wave and per SIMD-cycle from a pass with
        SQ_WAVES, SQ_INSTS_VALU and GRBM_GUI_ACTIVE

Every figure is printed as a markdown table row.  The source digest of the
library's kernels (`source_digest()`) is printed with the tables, so that a
committed summary can be matched to the sources it was measured on.
"""
import csv
import glob
import hashlib
import os
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
CUS, SIMDS, XCDS = 256, 4, 8


def source_digest():
    """sha256 over the library's kernel sources (csrc/*.hip, csrc/*.h)"""
    h = hashlib.sha256()
    for p in sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")) +
                    glob.glob(os.path.join(HERE, "csrc", "*.h"))):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def _rows(d, pattern):
    out = []
    for p in glob.glob(os.path.join(d, "**", pattern), recursive=True):
        with open(p) as f:
            out += list(csv.DictReader(f))
    return out


def short(name):
    """kernel name without the argument list"""
    name = name.split("(")[0]
    return name.replace("void ", "").replace("nyl::", "")


def trace(d):
    """{kernel: (launches, mean_us, min_us, total_ms)} from a kernel trace"""
    rows = _rows(d, "*kernel_trace.csv")
    acc = defaultdict(list)
    for r in rows:
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        grid = "%sx%sx%s" % (r.get("Grid_Size_X", "?"), r.get("Grid_Size_Y", "?"),
                             r.get("Grid_Size_Z", "?"))
        acc[(short(r["Kernel_Name"]), grid)].append(dur)
    return {k: (len(v), sum(v) / len(v), min(v), sum(v) / 1e3) for k, v in acc.items()}


def _counter(d, name):
    """{(kernel, grid): mean counter value per dispatch}"""
    rows = _rows(d, "*counter_collection.csv")
    acc = defaultdict(list)
    for r in rows:
        if r.get("Counter_Name") != name:
            continue
        grid = r.get("Grid_Size", "?")
        acc[(short(r["Kernel_Name"]), grid)].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main(argv):
    if not argv:
        print(__doc__)
        return 2
    what = argv[0]
    print("source digest:", source_digest())
    if what == "trace":
        for d in argv[1:]:
            t = trace(d)
            tot = sum(v[3] for v in t.values())
            print("\n%s (total kernel time %.3f ms)\n" % (d, tot))
            print("| kernel | grid | launches | mean µs | min µs | share |")
            print("|---|---|---|---|---|---|")
            for (k, grid), (n, mean, mn, ms) in sorted(t.items(), key=lambda kv: -kv[1][3]):
                if ms / tot < 0.005:
                    continue
                print("| `%s` | %s | %d | %.1f | %.1f | %.1f%% |" % (k, grid, n, mean, mn,
                                                                    100 * ms / tot))
        return 0
    if what == "traffic":
        fd, wd, cfd, cwd, nbytes = argv[1:6]
        nbytes = float(nbytes)
        f, w = _counter(fd, "FETCH_SIZE"), _counter(wd, "WRITE_SIZE")
        cf = max(_counter(cfd, "FETCH_SIZE").values())
        cw = max(_counter(cwd, "WRITE_SIZE").values())
        kf, kw = nbytes / (cf * 1024), nbytes / (cw * 1024)
        print("copy calibration: HBM read = FETCH_SIZE x 1024 x %.3f, "
              "HBM write = WRITE_SIZE x 1024 x %.3f\n" % (kf, kw))
        print("| kernel | grid | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB |")
        print("|---|---|---|---|---|---|")
        for key in sorted(f, key=lambda k: -(f[k] + w.get(k, 0.0))):
            fr, wr = f[key], w.get(key, 0.0)
            print("| `%s` | %s | %.0f | %.0f | %.1f | %.1f |" % (
                key[0], key[1], fr, wr, fr * 1024 * kf / 1e6, wr * 1024 * kw / 1e6))
        return 0
    if what == "valu":
        d = argv[1]
        waves, valu, gui = (_counter(d, "SQ_WAVES"), _counter(d, "SQ_INSTS_VALU"),
                            _counter(d, "GRBM_GUI_ACTIVE"))
        print("| kernel | grid | VALU instr. per wave | VALU instr. per SIMD-cycle |")
        print("|---|---|---|---|")
        for key in sorted(valu, key=lambda k: -valu[k]):
            wv = waves.get(key, 0.0)
            cyc = gui.get(key, 0.0) / XCDS
            rate = valu[key] / (CUS * SIMDS * cyc) if cyc > 0 else float("nan")
            print("| `%s` | %s | %.0f | %.3f |" % (key[0], key[1],
                                                  valu[key] / wv if wv else float("nan"), rate))
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
