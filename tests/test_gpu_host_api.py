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


def test_python_renderer_mirrors_reference_api():
    """K.Renderer (init / setup / frame / warm_up / cleanup): plain frames
    equal the C ABI's RGB render, SECONDARY_RAY frames the hybrid render."""
    import numpy as np
    import scenes as S
    hs = S.uniform_10k()
    r = K.Renderer()
    with pytest.raises(K.KdgrtError):
        r.setup(hs.tree, hs.g_inv)
    assert r.init()
    with pytest.raises(K.KdgrtError):
        r.frame(K.Camera(), 8, 8, None)
    r.setup(hs.tree, hs.g_inv)
    r.warm_up()
    cam = K.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 60.0, 1.0, 0.005)
    fb = r.dev.alloc(128 * 128 * 12)
    fps = r.frame(cam, 128, 128, fb)
    assert fps > 0
    got = fb.download(np.float32, (128, 128, 3))
    ci = K.camera_make(cam, 128, 128)
    assert np.array_equal(got, r.dev.render_image(r.scene, ci, 128, 128, fmt=K.OUT_RGB))
    r.setup(hs.tree, hs.g_inv, secondary={"spheres": K.demo_spheres()})
    assert r.frame(cam, 128, 128, fb) > 0
    got = fb.download(np.float32, (128, 128, 3))
    assert np.array_equal(got, r.dev.render_hybrid(r.scene, ci, 128, 128, spheres=K.demo_spheres()))
    fb.free()
    r.cleanup()
