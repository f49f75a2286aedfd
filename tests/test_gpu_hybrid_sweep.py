"""Seeded random sweep of the hybrid render (k_hybrid, SECONDARY_RAY,
cudaRenderer.cu:2217-2311) against the FP32 oracle.

Each case takes a random scene and view as in tests/test_gpu_sweep.py and
adds random secondary geometry: one to eight spheres of random material
(mirror, glass, diffuse), size, reflectivity, index of refraction and colour
around the scene, or a mesh of random triangles with one random material;
up to 16 bounces and a random stack depth.  The bar is that of the hybrid
parity tests: RGB within 1e-5 max abs of the oracle.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_sweep import _scene, _view

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

TOL = 1e-5
SEEDS = list(range(24))


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _spheres(rng, aabb):
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), float((hi - lo).max())
    n = int(rng.integers(1, 9))
    a = np.zeros(n, K.SECONDARY_SPHERE_DTYPE)
    for i in range(n):
        a[i] = (c + ext * rng.uniform(-0.4, 0.4, 3), ext * rng.uniform(0.03, 0.15),
                int(rng.integers(0, 3)), rng.uniform(0.3, 1.0), rng.uniform(1.0, 1.8), rng.uniform(0.2, 1.0, 3))
    return a


def _mesh(rng, aabb):
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), float((hi - lo).max())
    n = int(rng.integers(2, 13))
    centres = c + ext * rng.uniform(-0.35, 0.35, (n, 1, 3))
    tris = centres + ext * rng.uniform(-0.2, 0.2, (n, 3, 3))
    mat = (int(rng.integers(0, 3)), float(rng.uniform(0.3, 1.0)), float(rng.uniform(1.0, 1.8)),
           *[float(x) for x in rng.uniform(0.2, 1.0, 3)])
    return tris.reshape(-1, 9).astype(np.float32), mat


@pytest.mark.parametrize("seed", SEEDS)
def test_random_hybrid(dev, seed):
    rng = np.random.default_rng(9000 + seed)
    hs = _scene(rng)
    ci, w, h = _view(rng, hs.aabb)
    depth = int(rng.integers(2, 25))
    bounces = int(rng.integers(1, 17))
    sc = dev.scene(hs.tree, hs.g)
    if rng.random() < 0.7:
        sph = _spheres(rng, hs.aabb)
        print(f"[hybrid sweep {seed}] {len(hs.g)} Gaussians, {w}x{h}, {len(sph)} spheres, "
              f"{bounces} bounces, depth {depth}", flush=True)
        gpu = dev.render_hybrid(sc, ci, w, h, spheres=sph, max_bounces=bounces, depth=depth)
        ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h, spheres=sph, max_bounces=bounces,
                              depth=depth)
    else:
        tri, mat = _mesh(rng, hs.aabb)
        print(f"[hybrid sweep {seed}] {len(hs.g)} Gaussians, {w}x{h}, {len(tri)} triangles (material {mat[0]}), "
              f"{bounces} bounces, depth {depth}", flush=True)
        gpu = dev.render_hybrid(sc, ci, w, h, triangles=tri, mesh_material=mat[0], mesh_reflectivity=mat[1],
                                mesh_ior=mat[2], mesh_color=mat[3:], max_bounces=bounces, depth=depth)
        ref = O.render_hybrid(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h, triangles=tri, mesh=mat,
                              max_bounces=bounces, depth=depth)
    sc.free()
    err = np.abs(gpu.astype(np.float64) - ref["rgb"].astype(np.float64))
    print(f"  max err {err.max():.3g}, identical {float(np.mean(gpu == ref['rgb'])):.6f}, "
          f"rays with >1 segment {float(np.mean(ref['segments'] > 1)):.4f}", flush=True)
    assert err.max() <= TOL, f"seed {seed}: max |d rgb| = {err.max()}"
