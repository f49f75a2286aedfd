"""Hybrid render on the GPU (k_hybrid, renderKernelGaussian_Mesh_Hybrid,
cudaRenderer.cu:2217-2311) against the FP32 oracle: RGB within 1e-5 max abs
(observed: identical bits) over scenes with the reference's demo spheres,
mirror / glass / diffuse spheres, triangle meshes of each material, several
stack depths and bounce limits, and the full 1M-Gaussian 1920x1080 view.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _check(gpu, ref, label):
    err = np.abs(gpu.astype(np.float64) - ref["rgb"].astype(np.float64))
    same = float(np.mean(gpu == ref["rgb"]))
    multi = float(np.mean(ref["segments"] > 1))
    print(f"[{label}] max err {err.max():.3g}, identical {same:.6f}, rays with >1 segment {multi:.4f}", flush=True)
    assert err.max() <= TOL, f"{label}: max |d rgb| = {err.max()}"
    return multi


@pytest.mark.parametrize("depth", [12, 4])
def test_demo_spheres_uniform_10k(dev, depth):
    hs = S.uniform_10k()
    ci = S.cam_c0()
    sc = dev.scene(hs.tree, hs.g)
    sph = K.demo_spheres()
    print(f"[hybrid demo spheres] D={depth}", flush=True)
    gpu = dev.render_hybrid(sc, ci, 256, 256, spheres=sph, depth=depth)
    ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 256, 256, spheres=sph, depth=depth)
    assert _check(gpu, ref, f"demo spheres D={depth}") > 0.01
    sc.free()


@pytest.mark.parametrize("max_bounces", [5, 2, 1])
def test_spheres_clustered_100k(dev, max_bounces):
    hs = S.clustered(100000, seed=1002)
    ci = S.cam_orbit(hs.aabb, 320, 240)
    sc = dev.scene(hs.tree, hs.g)
    sph = S.spheres_in_view(hs.aabb)
    print(f"[hybrid spheres 100k] bounces={max_bounces}", flush=True)
    gpu = dev.render_hybrid(sc, ci, 320, 240, spheres=sph, max_bounces=max_bounces)
    ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 320, 240, spheres=sph,
                          max_bounces=max_bounces)
    multi = _check(gpu, ref, f"spheres 100k bounces={max_bounces}")
    assert (multi > 0.02) == (max_bounces > 1)
    sc.free()


@pytest.mark.parametrize("mesh", [(1, 0.9, 1.0, 0.8, 0.85, 0.9), (2, 1.0, 1.4, 0.9, 0.95, 1.0),
                                  (0, 0.0, 1.0, 0.3, 0.6, 0.2)])
def test_triangle_mesh_clustered_30k(dev, mesh):
    hs = S.clustered(30000, seed=1003)
    ci = S.cam_orbit(hs.aabb, 256, 192)
    sc = dev.scene(hs.tree, hs.g)
    tri = S.mesh_in_view(hs.aabb)
    print(f"[hybrid mesh] material {mesh[0]}", flush=True)
    gpu = dev.render_hybrid(sc, ci, 256, 192, triangles=tri, mesh_material=int(mesh[0]), mesh_reflectivity=mesh[1],
                            mesh_ior=mesh[2], mesh_color=mesh[3:])
    ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 256, 192, triangles=tri, mesh=mesh)
    multi = _check(gpu, ref, f"mesh material {mesh[0]}")
    assert (multi > 0.02) == (mesh[0] != 0)
    sc.free()


def test_no_secondary_equals_plain(dev):
    hs = S.uniform_10k()
    ci = S.cam_c0()
    sc = dev.scene(hs.tree, hs.g)
    print("[hybrid] no secondary primitives", flush=True)
    gpu = dev.render_hybrid(sc, ci, 256, 256, spheres=K.demo_spheres()[:0])
    plain = dev.render_image(sc, ci, 256, 256, fmt=K.OUT_RGB)
    assert np.array_equal(gpu, np.minimum(plain, np.float32(1)))
    sc.free()


def test_rejects_bad_secondary(dev):
    hs = S.uniform_10k()
    sc = dev.scene(hs.tree, hs.g)
    ci = S.cam_c0(16, 16)
    many = np.zeros(9, K.SECONDARY_SPHERE_DTYPE)
    with pytest.raises(K.KdgrtError):
        dev.render_hybrid(sc, ci, 16, 16, spheres=many)
    with pytest.raises(K.KdgrtError):
        dev.render_hybrid(sc, ci, 16, 16, spheres=K.demo_spheres(), max_bounces=0)
    sc.free()


def test_full_size_1m_1080p_hybrid(dev):
    """1M clustered Gaussians, one 1920x1080 view from radius 25 with three
    secondary spheres in view, every ray against the FP32 oracle."""
    hs = S.clustered(1_000_000)
    ci = S.cam_orbit(hs.aabb, 1920, 1080)
    sc = dev.scene(hs.tree, hs.g)
    sph = S.spheres_in_view(hs.aabb)
    print("[hybrid 1M 1080p] device render", flush=True)
    gpu = dev.render_hybrid(sc, ci, 1920, 1080, spheres=sph)
    print("[hybrid 1M 1080p] oracle render", flush=True)
    ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 1920, 1080, spheres=sph)
    assert _check(gpu, ref, "hybrid 1M 1080p") > 0.01
    sc.free()
