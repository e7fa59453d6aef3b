"""Builds libddx.so (HIP kernels for gfx950 + host C ABI) and the command-line tools, in-tree.

    python -m dds_gpu_amd.build          # or: from dds_gpu_amd.build import build; build()

Outputs (git-ignored, shipped with the tree):
    dds_gpu_amd/lib/libddx.so            the C ABI of include/ddx.h
    dds_gpu_amd/bin/dd_cpu, dd_gpu, dd_stream, dd_train, dd_pbn, pbn, dd_verify   command lines
    dds_gpu_amd/lib/libddx_sig.so        the same ABI built with 12-byte signature TT entries
    dds_gpu_amd/bin/sig/dd_cpu, dd_gpu, dd_stream   its command lines
                                         (reference: make ... NVFLAGS=-DDD_TT_SIG)
    dds_gpu_amd/lib/libddx_shared.so     the same ABI with one lock-free TT shared by all searches
    dds_gpu_amd/bin/shared/dd_cpu, dd_gpu, dd_stream   its command lines (reference: -DDD_SHARED_TT)
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
# Build variants: name -> (defines, tools). "" is the default exact-entry library.
VARIANTS = {"": ([], TOOLS), "sig": (["-DDDX_TT_SIG=1"], ["dd_cpu", "dd_gpu", "dd_stream"]),
            "shared": (["-DDDX_SHARED_TT=1"], ["dd_cpu", "dd_gpu", "dd_stream"])}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile_jobs(variant):
    defines = VARIANTS[variant][0]
    obj = os.path.join(OBJ, variant) if variant else OBJ
    os.makedirs(obj, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(obj, src + ".o")
        jobs.append([HIPCC] + DEVICE_FLAGS + COMMON + defines + ["-c", path, "-o", out])
    for src in HOST_SOURCES:
        out = os.path.join(obj, src + ".o")
        jobs.append([CXX] + COMMON + HOST_ARCH + defines + ["-c", os.path.join(CSRC, src), "-o", out])
    return jobs


def _link(variant, objs, verbose):
    suffix = "_" + variant if variant else ""
    lib = os.path.join(LIB, "libddx%s.so" % suffix)
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, lib)
    bindir = os.path.join(BIN, variant) if variant else BIN
    os.makedirs(bindir, exist_ok=True)
    rpath = "$ORIGIN/../../lib" if variant else "$ORIGIN/../lib"
    tools = []
    for t in VARIANTS[variant][1]:
        src = os.path.join(PKG, "cli", t + ".cpp")
        if not os.path.exists(src):
            continue
        tools.append([CXX] + COMMON + HOST_ARCH + VARIANTS[variant][0] +
                     [src, "-o", os.path.join(bindir, t), "-L" + LIB, "-lddx" + suffix,
                      "-Wl,-rpath," + rpath, "-lpthread"])
    return lib, tools


def build(verbose=False):
    """Compiles every HIP source for gfx950 and links libddx.so, its signature-entry and
    shared-TT variants (libddx_sig.so, libddx_shared.so), and the tools."""
    os.makedirs(LIB, exist_ok=True)
    jobs = {v: _compile_jobs(v) for v in VARIANTS}
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(lambda c: _run(c, verbose), [c for v in VARIANTS for c in jobs[v]]))
    tools = []
    for v in VARIANTS:
        lib, t = _link(v, [c[-1] for c in jobs[v]], verbose)
        tools += t
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(lambda c: _run(c, verbose), tools))
    return os.path.join(LIB, "libddx.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
