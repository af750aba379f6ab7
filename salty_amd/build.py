"""Builds the gfx950 shared library ``salty_amd/libs3dg.so`` in-tree with hipcc.

Every HIP source is compiled for gfx950 only.  Flags that matter for results:

* ``-ffp-contract=off``: no implicit FMA, so the ray/triangle, shading and
  tonemap arithmetic is one IEEE operation per source operation, as in the
  reference's /arch:AVX build.  Kernels use ``fmaf`` explicitly where they want
  a fused multiply-add (slab tests, which only cull).
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and square root.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU
  instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libs3dg.so")

SOURCES = ["bvh.cpp", "scene.cpp", "io.cpp", "trace.hip", "render.hip", "sceneset.hip", "post.hip", "cnn.hip",
           "dropin.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build(jobs: int = 8) -> str:
    """Compiles every source to an object in parallel, then links libs3dg.so."""
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # C++ drop-in programs: plain g++ against the C ABI, no HIP headers.
    host = os.path.join(HERE, "host")
    for src, exe in (("cornell_main.cpp", "s3d_cornell"), ("iface_demo.cpp", "s3d_iface_demo")):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", host, os.path.join(host, src),
              "-L", HERE, "-ls3dg", "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(host, exe)])
    return LIB


def _wait(item) -> None:
    cmd, proc = item
    out, _ = proc.communicate()
    out = "\n".join(line for line in out.splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


if __name__ == "__main__":
    print(build())
