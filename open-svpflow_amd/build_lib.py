"""Builds libsvpf.so (HIP kernels for gfx950 + C-ABI host code) in-tree.

    python open-svpflow_amd/build_lib.py

Every .hip / .cpp under csrc/ is compiled with hipcc for gfx950 and linked into
open-svpflow_amd/libsvpf.so.  Objects go to open-svpflow_amd/_build/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsvpf.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# VapourSynth plugin shells over libsvpf (one shared library per reference plugin)
PLUGINS = [("libsvpflow1_vs.so", "svp1.cpp"), ("libsvpflow2_vs.so", "svp2.cpp")]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-I", os.path.join(HERE, "..", "include")]
# Device code: gfx950 only; packed-FP32 instruction selection disabled.
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-x", "c++", "-c", src,
               "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    for name, src in PLUGINS:
        out = os.path.join(HERE, name)
        cmd = [CXX, "-O2", "-fPIC", "-std=c++17", "-Wall", "-fvisibility=hidden", "-I", os.path.join(HERE, "..", "include"),
               "-shared", "-o", out + ".tmp", os.path.join(HERE, "vsplugin", src), "-L", HERE, "-lsvpf",
               "-Wl,-rpath,$ORIGIN", "-Wl,--version-script=" + os.path.join(HERE, "vsplugin", "exports.map")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"plugin build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
