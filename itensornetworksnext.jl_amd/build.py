"""Compile the HIP engine for gfx950 into csrc/libbpmi355x.so (in-tree).

One translation unit (csrc/bp_engine.hip) with the kernels in its headers.  Packed FP32
instructions are disabled for every device compile (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(CSRC, "libbpmi355x.so")
SOURCES = ["bp_engine.hip"]
HEADERS = ["common.h", "generic.h", "complex.h", "vertex.h", "vertex16.h", "vertex12.h",
           "vertex64.h", "vertex64l.h", "vertex4.h", "adjoint.h", "vertex_adjoint.h", "contract.h", "gate.h"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP engine cannot be built")


def compile_command(out: str = LIB):
    return [hipcc(), "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-function", "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


def stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    deps = SOURCES + HEADERS
    inc = os.path.join(os.path.dirname(HERE), "include", "bp_mi355x.h")
    paths = [os.path.join(CSRC, d) for d in deps] + [inc]
    return any(os.path.getmtime(p) > t for p in paths if os.path.exists(p))


def build(force: bool = False, verbose: bool = False) -> str:
    if not force and not stale():
        return LIB
    tmp = LIB + ".tmp"
    proc = subprocess.run(compile_command(tmp), capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + proc.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print(proc.stderr)
    return LIB


if __name__ == "__main__":
    print(build(force=True, verbose=True))
