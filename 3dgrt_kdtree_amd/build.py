"""Builds the in-tree shared library lib/libkdgrt.so for gfx950.

Host C++ (builder, I/O, serving loop) is compiled by g++ with OpenMP; the
HIP translation units by hipcc for --offload-arch=gfx950.  Everything is
compiled with -ffp-contract=off: the render path's float expressions must be
evaluated operation by operation on host and device alike.
"""
import os
import shutil
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libkdgrt.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["host_core.cpp", "sah_builder.cpp", "ingest_stats.cpp"]
HIP_SOURCES = ["render.hip", "autotune.hip", "server.hip", "host_api.hip", "viewer.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden", "-I" + INC, "-I" + CSRC]
HOST_FLAGS = COMMON + ["-fopenmp", "-fno-math-errno", "-Wall", "-Wno-unknown-pragmas"]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    "-fno-gpu-rdc",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # the iterative ILP instruction scheduler: the render kernels are bound
    # by vector-ALU and texture-path issue together, and this order of their
    # instructions measured faster everywhere (1080p view -1.5 %, 64-view
    # batch -4.5 %, 4M rays -7 %, hybrid -2 %, serving burst -0.5 %; the
    # default and max-ilp / max-memory-clause / iterative-minreg / legacy
    # strategies were slower or level), results bit-identical
    "-mllvm", "--amdgpu-sched-strategy=iterative-ilp",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj])
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-fopenmp",
        "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
