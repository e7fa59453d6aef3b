"""Builds libddx.so (HIP kernels for gfx950 + host C ABI) and the command-line tools, in-tree.

    python -m dds_gpu_amd.build          # or: from dds_gpu_amd.build import build; build()

Outputs (git-ignored, shipped with the tree):
    dds_gpu_amd/lib/libddx.so            the C ABI of include/ddx.h
    dds_gpu_amd/bin/dd_cpu, dd_gpu, dd_stream, dd_train, dd_pbn, pbn, dd_verify   command lines
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dds_gpu_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "c++")
ARCH = "gfx950"
# Host sources and the tools are plain C++ (g++); only .hip files go through hipcc.
# Packed FP32 instructions are disabled for all device code (no float math relies on them).
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unknown-pragmas", "-I" + INC, "-I" + CSRC]
HOST_ARCH = ["-march=x86-64-v3"]

HIP_SOURCES = ["ddx_gpu.hip", "ddx_mlp.hip", "ddx_stream.hip"]
HOST_SOURCES = ["ddx_host.cpp"]
TOOLS = ["dd_cpu", "dd_gpu", "dd_stream", "dd_train", "dd_pbn", "pbn", "dd_verify"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _objects(verbose):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC] + DEVICE_FLAGS + COMMON + ["-c", path, "-o", out])
    for src in HOST_SOURCES:
        out = os.path.join(OBJ, src + ".o")
        jobs.append([CXX] + COMMON + HOST_ARCH + ["-c", os.path.join(CSRC, src), "-o", out])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    return [j[-1] for j in jobs]


def build(verbose=False):
    """Compiles every HIP source for gfx950 and links libddx.so and the tools."""
    objs = _objects(verbose)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    lib = os.path.join(LIB, "libddx.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, lib)
    tools = []
    for t in TOOLS:
        src = os.path.join(PKG, "cli", t + ".cpp")
        if not os.path.exists(src):
            continue
        out = os.path.join(BIN, t)
        tools.append([CXX] + COMMON + HOST_ARCH + [src, "-o", out, "-L" + LIB, "-lddx",
                                                     "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"])
    if tools:
        with ThreadPoolExecutor(max_workers=min(8, len(tools))) as ex:
            list(ex.map(lambda c: _run(c, verbose), tools))
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
