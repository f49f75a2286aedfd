"""Seeded test scenes shared by the CPU and GPU suites.

Scene families (seed = 1000 + configuration index, as in the measurement
plan): the 10k uniform-cube scene rendered at 256x256 from (0,0,3); the
clustered family at reduced counts for parity runs; hand-built trees.
"""
import functools
import importlib

import numpy as np

K = importlib.import_module("3dgrt_kdtree_amd")


class HostScene:
    def __init__(self, g, tree):
        self.g = g                       # activated, non-inverted scales
        self.g_inv = K.invert_scales(g)  # the renderer's input (inverted scales)
        self.tree = tree
        self.nodes = tree.nodes
        self.offsets = tree.offsets
        self.aabb = tree.aabb


@functools.lru_cache(maxsize=None)
def uniform_10k():
    """10k Gaussians, mu ~ U[-1,1]^3, s = exp(N(ln .01,.5)), o = sigmoid(N(0,1))."""
    g = K.generate_gaussians(0, 10000, 1000)
    return HostScene(g, K.Tree.build(g))


@functools.lru_cache(maxsize=None)
def clustered(n, seed=1001):
    g = K.generate_gaussians(1, n, seed)
    return HostScene(g, K.Tree.build(g))


@functools.lru_cache(maxsize=None)
def anisotropic(n=3000, seed=41):
    """Needles and pancakes: one axis of each Gaussian 1, 10 or 100 times
    the others, random rotations, in the unit cube."""
    rng = np.random.default_rng(seed)
    g = np.zeros(n, dtype=K.GAUSSIAN_DTYPE)
    g["pos"] = rng.uniform(-1, 1, (n, 3))
    s = np.exp(rng.normal(np.log(0.01), 0.6, (n, 3)))
    s[np.arange(n), rng.integers(0, 3, n)] *= rng.choice([1.0, 10.0, 100.0], n)
    g["scale"] = np.minimum(s, 0.5)
    q = rng.normal(size=(n, 4))
    g["rot"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    g["opacity"] = 1.0 / (1.0 + np.exp(-rng.normal(0, 1.5, n)))
    g["k_scale"] = 1.0
    g["f_dc"] = (rng.random((n, 3)) - 0.5) / 0.28209479177387814
    return HostScene(g, K.Tree.build(g))


def cam_c0(w=256, h=256):
    """Eye (0,0,3), n = (0,0,1), u = (1,0,0), v = (0,1,0), fovy 60, near 0.005."""
    c = K.Camera()
    c.pos[:] = (0.0, 0.0, 3.0)
    c.u[:] = (1.0, 0.0, 0.0)
    c.v[:] = (0.0, 1.0, 0.0)
    c.n[:] = (0.0, 0.0, 1.0)
    c.fovy_deg, c.aspect, c.near_c = 60.0, w / h, 0.005
    return K.camera_make(c, w, h)


def cam_orbit(aabb, w, h, radius=25.0, direction=(0.48, 0.36, 0.8), fovy=60.0):
    centre = 0.5 * (np.asarray(aabb[:3], np.float64) + np.asarray(aabb[3:], np.float64))
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    cam = K.look_at(centre + radius * d, centre, fovy, w / h, 0.005)
    return K.camera_make(cam, w, h)


def random_rays(aabb, n, seed):
    """Origins uniform in the scene box, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(aabb[:3]), np.asarray(aabb[3:])
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def single_leaf_scene(n=12, seed=7, opacity=0.3):
    """At most 16 kept Gaussians: the builder makes a single leaf
    (Kd-treeConstructor.cpp:1503), so the render equals brute force.
    Gaussians are strung along the z axis so one ray meets all of them."""
    rng = np.random.default_rng(seed)
    g = np.zeros(n, dtype=K.GAUSSIAN_DTYPE)
    g["pos"][:, 0] = rng.normal(0, 0.01, n)
    g["pos"][:, 1] = rng.normal(0, 0.01, n)
    g["pos"][:, 2] = np.linspace(-1, 1, n) + rng.normal(0, 0.01, n)
    g["scale"][:] = np.exp(rng.normal(np.log(0.05), 0.3, (n, 3)))
    q = rng.normal(size=(n, 4))
    g["rot"][:] = q / np.linalg.norm(q, axis=1, keepdims=True)
    g["opacity"][:] = opacity
    g["k_scale"][:] = 1.0
    g["f_dc"][:] = rng.uniform(-1, 1, (n, 3))
    g["f_rest"][:] = rng.normal(0, 0.2, (n, 45))
    return HostScene(g, K.Tree.build(g))


def spheres_in_view(aabb):
    """Three secondary spheres around the scene centre, sized to the box: a
    glass sphere (ior 1.3), a tinted mirror (reflectivity 0.85) and a
    diffuse one."""
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), float((hi - lo).max())
    a = np.zeros(3, K.SECONDARY_SPHERE_DTYPE)
    a[0] = (c + 0.10 * ext * np.array([1.0, 0.3, 0.2]), 0.08 * ext, K.MATERIAL_GLASS, 1.0, 1.3, (0.95, 0.9, 0.85))
    a[1] = (c - 0.12 * ext * np.array([1.0, -0.2, 0.1]), 0.10 * ext, K.MATERIAL_MIRROR, 0.85, 1.0, (0.6, 0.7, 0.9))
    a[2] = (c + 0.20 * ext * np.array([0.0, 1.0, 0.0]), 0.04 * ext, K.MATERIAL_DIFFUSE, 0.0, 1.0, (0.8, 0.2, 0.1))
    return a


def mesh_in_view(aabb):
    """A mirror quad (two triangles) through the scene centre, tilted, and
    a tetrahedron of four triangles beside it."""
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), float((hi - lo).max())
    s = 0.3 * ext
    q = [c + s * np.array(v) for v in ([-1, -1, 0.3], [1, -1, -0.3], [1, 1, -0.3], [-1, 1, 0.3])]
    t0 = c + 0.25 * ext * np.array([0.6, 0.4, 0.5])
    tv = [t0 + 0.1 * ext * np.array(v) for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])]
    tris = [q[0], q[1], q[2], q[0], q[2], q[3],
            tv[0], tv[1], tv[2], tv[0], tv[3], tv[1], tv[0], tv[2], tv[3], tv[1], tv[3], tv[2]]
    return np.asarray(tris, np.float32).reshape(-1, 9)
