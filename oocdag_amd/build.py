"""Builds liboocdag.so (HIP for gfx950) and the oocdag CLI in-tree.

    python -m oocdag_amd.build          # or oocdag_amd.build.build()

hipcc compiles every .hip source for gfx950 only; the host side uses
OpenMP for the tile packing between column-major A and the pinned slots.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liboocdag.so")
CLI = os.path.join(HERE, "oocdag")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LLVM_LIB = "/opt/rocm/lib/llvm/lib"

SOURCES = ["gemm.hip", "gemm_lp.hip", "factor.hip", "solve.hip", "mixed.hip", "batched.hip", "serve.hip", "driver.hip"]

HIPFLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fopenmp", "-o", tmp, *objs,
          f"-Wl,-rpath,{ROCM_LLVM_LIB}", f"-L{ROCM_LLVM_LIB}"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "oocdag_main.cpp")
    if os.path.exists(cli_src):
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src,
              "-o", CLI + ".tmp", f"-L{HERE}", "-loocdag", "-Wl,-rpath,$ORIGIN",
              f"-Wl,-rpath,{ROCM_LLVM_LIB}"])
        os.replace(CLI + ".tmp", CLI)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
