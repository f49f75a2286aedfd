"""C ABI checks that need no GPU: every declared symbol is exported, error
codes and messages instead of exits, argument validation."""
import ctypes as C
import importlib
import os
import re

import pytest

K = importlib.import_module("3dgrt_kdtree_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        names += re.findall(r"KDGRT_API\s+[\w\s\*]+?\b(kdgrt_\w+)\s*\(", txt)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) > 40
    lib = C.CDLL(K.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_reference_named_cpp_api_is_exported():
    """Every function of include/kdgrt_reference_api.hpp is a default-
    visibility symbol of the library (demangled names from nm -DC)."""
    import subprocess
    out = subprocess.run(["nm", "-DC", K.LIB_PATH], stdout=subprocess.PIPE, text=True, check=True).stdout
    txt = open(os.path.join(ROOT, "include", "kdgrt_reference_api.hpp")).read()
    declared = set(re.findall(r"KDGRT_REF_API\s+[\w:<>\s\*&]+?\b(\w+)\s*\(", txt))
    assert {"initCuda", "renderGaussianWithCudaSetup", "renderGaussianWithCudaFrame", "warmUp",
            "cleanupCudaResources", "setSecondaryPrim", "loadSecondaryMesh"} <= declared
    exported = set(re.findall(r" T kdgrt_ref::(\w+)\(", out))
    assert declared <= exported, declared - exported
    # both set-up overloads (plain and SECONDARY_RAY)
    assert out.count("kdgrt_ref::renderGaussianWithCudaSetup(") == 2


def test_library_is_built_for_gfx950():
    data = open(K.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_errors_are_codes_not_exits():
    L = K.lib()
    g = (C.c_char * 256)()
    assert L.kdgrt_generate_gaussians(C.c_int(7), C.c_uint64(1), C.c_uint64(0), g) == 1
    assert b"kind" in L.kdgrt_last_error()
    ci = K.CameraInfo()
    assert L.kdgrt_camera_make(None, 10, 10, C.byref(ci)) == 1
    cam = K.Camera()
    assert L.kdgrt_camera_make(C.byref(cam), 0, 10, C.byref(ci)) == 1
    h = C.c_void_p()
    assert L.kdgrt_tree_build(None, C.c_uint64(0), None, C.byref(h)) == 1
    assert L.kdgrt_kdt_read(b"/nonexistent/x.kdt", C.byref(h)) == 4
    assert L.kdgrt_scene_set_tile_schedule(None, C.c_int(1)) == 1
    assert b"null scene" in L.kdgrt_last_error()


def test_init_without_device_reports_no_device():
    n = C.c_int()
    K.check(K.lib().kdgrt_device_count(C.byref(n)))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    h = C.c_void_p()
    assert K.lib().kdgrt_init(0, C.byref(h)) == 8


def test_all_culled_scene_is_an_error():
    g = K.generate_gaussians(0, 10, 1)
    g["opacity"][:] = 0.001
    with pytest.raises(K.KdgrtError):
        K.Tree.build(g)


def test_python_renderer_call_order_errors():
    """K.Renderer reports the reference's call-order errors as exceptions
    (setup before init, frame before setup) without touching a device."""
    r = K.Renderer()
    with pytest.raises(K.KdgrtError):
        r.setup(None, [])
    with pytest.raises(K.KdgrtError):
        r.frame(K.Camera(), 8, 8, None)
    with pytest.raises(K.KdgrtError):
        r.warm_up()
    r.cleanup()
