"""Build step of the package: compiles the HIP library for gfx950 in-tree.

``python -m yelmo_amd.build`` (or ``__graft_entry__.build()``) produces
``yelmo_amd/libygpu.so`` from ``yelmo_amd/csrc/*.hip`` with hipcc.  The
library is loaded by ``yelmo_amd._lib``; nothing is compiled at import time.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libygpu.so"
SOURCES = ["ygpu_api.hip", "ygpu_serve.hip", "ygpu_tune.hip", "ygpu_fields.hip",
           "ygpu_krylov.hip", "ygpu_krylov_1024.hip", "ygpu_krylov_group.hip", "ygpu_diva.hip",
           "ygpu_spmv.hip", "ygpu_noalloc.hip", "ygpu_scalar.hip", "ygpu_inputs.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions: hipcc must not form v_pk_* FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> Path:
    objs = []
    jobs = []
    for src in SOURCES:
        obj = CSRC / (src + ".o")
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", str(CSRC / src), "-o", str(obj)]
        jobs.append((cmd, obj))
    procs = [(subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), o, c)
             for c, o in jobs]
    for p, o, c in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed for {o.name}:\n" + "\n".join(lines))
        if verbose and lines:
            print("\n".join(lines))
        objs.append(str(o))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
