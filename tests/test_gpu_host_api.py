"""The reference-named C++ host API (initCuda / renderGaussianWithCudaSetup /
renderGaussianWithCudaFrame / warmUp / cleanupCudaResources) driven from C++
like the reference's render-test mode; its frame must equal the C ABI render."""
import importlib
import os
import subprocess

import pytest

K = importlib.import_module("3dgrt_kdtree_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_reference_named_host_api(tmp_path):
    exe = str(tmp_path / "host_api_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "host_api_driver.cpp"), "-o", exe,
                    "-L" + os.path.dirname(K.LIB_PATH), "-lkdgrt", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + os.path.dirname(K.LIB_PATH) + ":/opt/rocm/lib"], check=True)
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "host_api ok=1" in r.stdout
