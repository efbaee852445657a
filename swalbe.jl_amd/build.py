"""Builds libswalbe_hip.so for gfx950 with hipcc (no GPU needed).

    python swalbe.jl_amd/build.py [--jobs N]

Every translation unit is compiled with -ffp-contract=off (FP64 parity with
the reference's unfused Float64 arithmetic) and without packed-FP32
instructions; the objects are linked into swalbe.jl_amd/lib/libswalbe_hip.so.
The access-pattern probe bench.py measures its ceiling with is a separate
library, swalbe.jl_amd/lib/libswalbe_probe.so: it is no part of the
library's interface.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libswalbe_hip.so")
SOURCES = ["api.hip", "ops.hip", "step.hip", "step_f32.hip", "batch.hip", "serve.hip",
           "d1q3.hip", "measure.hip"]
PROBE_SOURCES = ["probe.hip"]
PROBE_LIB = os.path.join(LIBDIR, "libswalbe_probe.so")
ARCH = os.environ.get("SWALBE_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def compile_one(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def link(objs, out: str) -> None:
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build(jobs: int = 4) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES + PROBE_SOURCES))
    link(objs[:len(SOURCES)], LIB)
    link(objs[len(SOURCES):], PROBE_LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        print(build(a.jobs))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
