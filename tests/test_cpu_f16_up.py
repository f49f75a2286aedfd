"""The host's rounding of bound-sphere radii and dilations up to half
precision (csrc/f16_up.hpp), checked against every float.

The bound-sphere pre-test is conservative only if its stored radius and
dilation are not below the values derived for them; the host forms of
hip_fp16's directed conversions round to nearest, so the library rounds from
the bits and this test compares that with the smallest half >= x found by
search, for all positive finite floats up to 65504 and the special inputs.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "3dgrt_kdtree_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_f16_round_up_every_float(tmp_path):
    exe = str(tmp_path / "f16_up_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", CSRC,
                    os.path.join(ROOT, "tests", "cpp", "f16_up_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
