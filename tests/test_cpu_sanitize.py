"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

1. The host sources (proxy meshing, the OpenMP SAH builder, .kdt / PLY / OBJ /
   JSON / PNG / statistics I/O) built with g++ -fsanitize=address,undefined
   and driven by tests/cpp/sanitize_driver.cpp: valid scenes at several
   thread counts, corrupt and fuzzed files.  Leak checking is on.
2. The whole library with host-side sanitizers (clang; the device code is
   not instrumented), loaded into a child Python with the clang ASan
   runtime preloaded, running the serving-loop scheduler on the modelled
   device (tests/test_cpu_server.py) and the C ABI checks.

Any sanitizer report aborts the child process and fails the test.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "3dgrt_kdtree_amd", "csrc")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = ["host_core.cpp", "sah_builder.cpp", "ingest_stats.cpp"]
HIP = ["render.hip", "autotune.hip", "server.hip", "host_api.hip", "viewer.hip"]


def _run_all(cmds, cwd):
    procs = [subprocess.Popen(c, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for c in cmds]
    for c, p in zip(cmds, procs):
        out, _ = p.communicate()
        assert p.returncode == 0, " ".join(c) + "\n" + out[-4000:]


def _clang_asan_runtime():
    clang = os.path.join(ROCM, "llvm", "bin", "clang++")
    if not os.path.exists(clang):
        return None, None
    rt = subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"], stdout=subprocess.PIPE,
                        text=True).stdout.strip()
    return clang, (rt if os.path.isabs(rt) and os.path.exists(rt) else None)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_ubsan(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
    flags = ["-O1", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-I" + INC, "-I" + CSRC] + san
    _run_all([["g++"] + flags + ["-c", os.path.join(CSRC, s), "-o", s + ".o"] for s in HOST], str(tmp_path))
    _run_all([["g++"] + flags + [os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp")] +
              [s + ".o" for s in HOST] + ["-o", "driver"]], str(tmp_path))
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="8")
    r = subprocess.run([str(tmp_path / "driver"), str(scratch)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-6000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-6000:]
    assert "0 failed expectation" in r.stdout


def test_server_scheduler_under_asan_ubsan(tmp_path):
    clang, rt = _clang_asan_runtime()
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    if clang is None or rt is None or not os.path.exists(hipcc):
        pytest.skip("clang ASan runtime or hipcc not available")
    common = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + INC, "-I" + CSRC, "-Wno-unused-result",
              "-Wno-unused-value"]
    host_san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g"]
    cmds = [[clang] + common + ["-fopenmp"] + host_san + ["-c", os.path.join(CSRC, s), "-o", s + ".o"] for s in HOST]
    # sanitizers on the host pass only: each -fsanitize= follows -Xarch_host
    hip_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-Xarch_host", "-fno-omit-frame-pointer"]
    cmds += [[hipcc] + common + ["--offload-arch=gfx950", "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang",
                                  "-packed-fp32-ops"] + hip_san + ["-c", os.path.join(CSRC, s), "-o", s + ".o"]
             for s in HIP]
    _run_all(cmds, str(tmp_path))
    lib = str(tmp_path / "libkdgrt_san.so")
    _run_all([[clang, "-shared", "-fsanitize=address,undefined", "-shared-libsan", "-fopenmp", "-o", lib] +
              [s + ".o" for s in HOST + HIP] + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                                                 "-Wl,-rpath," + os.path.join(ROCM, "lib")]], str(tmp_path))
    env = dict(os.environ, KDGRT_LIB_PATH=lib, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_cpu_server.py"), os.path.join(ROOT, "tests", "test_cpu_abi.py")],
                       cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-6000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-6000:]
