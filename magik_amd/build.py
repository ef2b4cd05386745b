"""Build recipe of libmagik_amd.so (hipcc, gfx950), run in-tree.

    python -m magik_amd.build          # or __graft_entry__.build()

Generates csrc/cie1931_table.inc from data/cie1931_2deg_1nm.csv, then
compiles csrc/*.hip (one object each, in parallel, under magik_amd/_build/) and
links magik_amd/libmagik_amd.so.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmagik_amd.so")
SOURCES = ["mgk_kernels.hip", "mgk_api.hip", "mgk_mlp.hip", "mgk_surrogate.hip", "mgk_surrogate_frame.hip", "mgk_display.hip",
           "mgk_manager.hip", "mgk_mesh.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++20", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-fvisibility=hidden",
    # packed FP32 (v_pk_*_f32) stays off: forms of it have produced silently wrong results
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def gen_table() -> str:
    src = os.path.join(HERE, "data", "cie1931_2deg_1nm.csv")
    rows = []
    with open(src) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            w, x, y, z = line.strip().split(",")
            rows.append((int(w), x, y, z))
    assert len(rows) == 471 and rows[0][0] == 360 and rows[-1][0] == 830
    body = ",\n".join(f"    {{{x}f, {y}f, {z}f}}" for _, x, y, z in rows)
    text = ("// Generated by magik_amd/build.py from data/cie1931_2deg_1nm.csv; do not edit.\n"
            "static const float kCie1931[471][3] = {\n" + body + "\n};\n")
    path = os.path.join(CSRC, "cie1931_table.inc")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if err.strip() and verbose:
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> str:
    """Compiles every source to an object (in parallel), then links the library."""
    from concurrent.futures import ThreadPoolExecutor
    gen_table()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in SOURCES]
    jobs = [[HIPCC, *compile_flags, *inc, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    _run([HIPCC, *FLAGS, *objs, "-o", OUT + ".tmp"], verbose)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
