"""Edge cases of the device path against the FP32 oracle, to the same bar as
the parity tests (counters, composited-ID hashes bit-exact; RGB and T within
1e-5, observed identical):

* image sizes that leave partial 8x8 tiles, down to one pixel;
* a camera inside the scene box (rays start inside, t_near = 1e-2 clips);
* rays with zero direction components (the reciprocal is +-inf) starting
  exactly on split planes (0 * inf = NaN split distances, where every
  comparison is false), on the box faces and outside the box;
* rays that miss the scene box entirely.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _same(gpu, ref, label):
    assert np.array_equal(gpu["counters"], ref["counters"]), f"{label}: counters differ"
    assert np.array_equal(gpu["hash"], ref["hash"]), f"{label}: composited-ID hash differs"
    err = np.abs(gpu["rgbt"].astype(np.float64) - ref["rgbt"].astype(np.float64)).max()
    assert err <= 1e-5, f"{label}: max |d rgbt| = {err}"
    print(f"[{label}] max err {err:.3g}, identical {np.mean(gpu['rgbt'] == ref['rgbt']):.6f}", flush=True)


@pytest.mark.parametrize("wh", [(1, 1), (13, 7), (65, 33), (8, 200), (300, 9)])
def test_partial_tiles(dev, wh):
    w, h = wh
    hs = S.uniform_10k()
    ci = S.cam_c0(w, h)
    sc = dev.scene(hs.tree, hs.g)
    print(f"[edge] {w}x{h}", flush=True)
    _same(dev.render_debug(sc, ci, w, h), O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h), f"{w}x{h}")
    sc.free()


def test_camera_inside_the_box(dev):
    hs = S.clustered(30000, seed=1003)
    lo, hi = np.asarray(hs.aabb[:3]), np.asarray(hs.aabb[3:])
    c = 0.5 * (lo + hi)
    cam = K.look_at(c, c + np.array([0.3, -0.2, 1.0]), 90.0, 4 / 3, 0.005)
    ci = K.camera_make(cam, 160, 120)
    sc = dev.scene(hs.tree, hs.g)
    print("[edge] camera inside the box", flush=True)
    gpu = dev.render_debug(sc, ci, 160, 120)
    _same(gpu, O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 160, 120), "inside")
    assert (gpu["counters"][..., 1] > 0).mean() > 0.99  # every ray starts in some leaf
    sc.free()


@pytest.mark.parametrize("eye", [(0.0, 0.0, 1.6), (0.9, 0.7, 1.4), (4.0, 3.0, 6.0), (0.2, -0.1, 0.3)])
def test_anisotropic_gaussians_view_cone(dev, eye):
    """The view pre-test's cone thresholds bound each Gaussian by its
    ellipsoid projected across the view direction: needle and pancake
    Gaussians seen from near (inside the cloud), middle and far eyes render
    exactly as the oracle's exact tests do."""
    hs = S.anisotropic()
    cam = K.look_at(np.asarray(eye), np.zeros(3), 70.0, 4 / 3, 0.005)
    ci = K.camera_make(cam, 200, 150)
    sc = dev.scene(hs.tree, hs.g)
    print(f"[edge] anisotropic Gaussians from {eye}", flush=True)
    gpu = dev.render_debug(sc, ci, 200, 150)
    ref = O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, 200, 150)
    _same(gpu, ref, f"anisotropic {eye}")
    assert ref["counters"][..., 4].sum() > 1000  # many blends: the bound is exercised
    img = dev.render_image(sc, ci, 200, 150)  # the plain kernel, tile schedule included
    assert np.array_equal(img.view(np.uint32), ref["rgbt"].reshape(150, 200, 4).view(np.uint32))
    sc.free()


def _axis_rays(hs, n_per, seed):
    """Rays with zero direction components from split planes, box faces and
    outside the box."""
    rng = np.random.default_rng(seed)
    nodes = np.asarray(hs.nodes)
    inner = nodes[(nodes["x"] & 7) != 3]
    splits = np.ascontiguousarray(inner["y"]).view(np.float32)
    axes = inner["x"] & 3
    lo, hi = np.asarray(hs.aabb[:3], np.float32), np.asarray(hs.aabb[3:], np.float32)
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (0, 0.6, 0.8), (0.6, 0, -0.8),
            (-0.8, 0.6, 0)]
    o, d = [], []
    for _ in range(n_per):
        k = int(rng.integers(0, len(inner)))
        a = int(axes[k])
        p = (lo + (hi - lo) * rng.random(3)).astype(np.float32)
        p[a] = splits[k]  # exactly on a split plane
        for dv in dirs:
            o.append(p.copy())
            d.append(dv)
        q = p.copy()
        q[int(rng.integers(0, 3))] = lo[0] if rng.random() < 0.5 else hi[0]  # on / near a box face
        o.append(q)
        d.append(dirs[int(rng.integers(0, len(dirs)))])
        r = (hi + (hi - lo) * rng.random(3)).astype(np.float32)  # outside, pointing away or along
        o.append(r)
        d.append(dirs[int(rng.integers(0, len(dirs)))])
    return np.asarray(o, np.float32), np.asarray(d, np.float32)


def test_zero_direction_components_on_split_planes(dev):
    hs = S.clustered(30000, seed=1003)
    o, d = _axis_rays(hs, 400, 17)
    sc = dev.scene(hs.tree, hs.g)
    print(f"[edge] {len(o)} axis-aligned rays on split planes", flush=True)
    gpu = dev.render_rays(sc, o, d, debug=True)
    ref = O.render_rays(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, o, d)
    _same(gpu, ref, "axis rays")
    assert (ref["counters"][:, 4] > 0).sum() > 100  # some of them blend
    sc.free()


def test_rays_missing_the_box(dev):
    hs = S.uniform_10k()
    n = 256
    o = np.tile(np.float32([5, 5, 5]), (n, 1))
    d = np.tile(np.float32([1, 0, 0]), (n, 1))
    d[::2] = (0, 0, 1)
    sc = dev.scene(hs.tree, hs.g)
    gpu = dev.render_rays(sc, o, d, debug=True)
    ref = O.render_rays(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, o, d)
    _same(gpu, ref, "missing rays")
    assert (gpu["rgbt"] == np.float32([0, 0, 0, 1])).all() and (gpu["counters"] == 0).all()
    sc.free()


@pytest.mark.parametrize("n", [0, 1, 37, 1024, 1024 * 5 + 37, 256 * 1024 + 1023])
def test_copy_kernel_exact(dev, n):
    """k_copy (the HBM-ceiling and counter-calibration kernel) copies every
    16-B word: whole chunks of 4 x 256 words and the word-by-word remainder."""
    import ctypes as C
    rng = np.random.default_rng(n)
    src_h = rng.integers(0, 2**32, size=(max(n, 1), 4), dtype=np.uint32)
    src = dev.upload(src_h)
    dst = dev.upload(np.zeros((max(n, 1) + 1, 4), np.uint32))  # one guard word past the end
    K.check(K.lib().kdgrt_copy_f32x4(dev.h, dst.ptr, src.ptr, C.c_uint64(n), None))
    dev.sync()
    got = dst.download(np.uint32, (max(n, 1) + 1, 4))
    assert np.array_equal(got[:n], src_h[:n]), "copied words differ"
    assert not got[n:].any(), "words past the copy were written"
    src.free()
    dst.free()
