"""Builds libh2eri_amd.so in-tree.

Device code (.hip) is compiled with hipcc for gfx950 only; host C++ is
compiled with g++/libgomp so the library shares PyTorch's OpenMP runtime.
Every call recompiles (no stale objects): sources are few and small.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "h2eri")
LIB = os.path.join(OUT_DIR, "libh2eri_amd.so")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_SOURCES = ["common.cpp", "eri.cpp", "idqr.cpp", "shellops.cpp", "partition.cpp", "build_h2.cpp", "rep_io.cpp"]
DEVICE_SOURCES = ["device.hip", "exchange.hip", "jit.hip"]

HOST_FLAGS = ["-O3", "-fPIC", "-fopenmp", "-std=c++17", "-march=x86-64-v3", "-Wall", "-Wno-unused-function"]
DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("compile failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + HOST_FLAGS + ["-I", INC, "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    for s in DEVICE_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + DEVICE_FLAGS + ["-I", INC, "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
