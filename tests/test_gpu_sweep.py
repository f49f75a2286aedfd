"""Seeded random sweep of the render path against the FP32 oracle.

Each case draws a scene, a view and a ray list from one seed:

* 500 to 20,000 Gaussians, uniform in a box or in a few clusters; per-axis
  scales log-normal, one axis of some Gaussians up to 30 times the others;
  random rotations; opacities sigmoid(N(0, 2)); half the scenes with
  non-zero degree 1-3 SH coefficients;
* a pinhole view from outside or inside the box, field of view 20 to 100
  degrees, 48 to 160 pixels a side, short-stack depth 2 to 24;
* 3,000 rays with origins inside and around the box and random directions.

The debug kernel must equal the oracle bit for bit (all eight counters, the
composited-ID hash and sequence, RGB and transmittance), the plain view
kernel must write the same image, and the ray list (plain and debug) must
equal the oracle's ray-list render.  The view and the ray list then run
together as one persistent batch under a random tuning (tile shape, hit
buffer 4, 7, 8 or 16 with its overflow re-pass, workgroup size, occupancy,
with or without graph replay), whose results must equal the oracle's too.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

SEEDS = list(range(48))


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def stream(dev):
    s = dev.stream()  # graph replay of a batch runs on an explicit stream
    yield s
    K.check(K.lib().kdgrt_stream_destroy(dev.h, s))


def _scene(rng):
    n = int(rng.integers(500, 20001))
    g = np.zeros(n, dtype=K.GAUSSIAN_DTYPE)
    half = float(rng.uniform(0.5, 5.0))
    if rng.random() < 0.5:
        g["pos"] = rng.uniform(-half, half, (n, 3))
    else:
        k = int(rng.integers(2, 9))
        centres = rng.uniform(-half, half, (k, 3))
        g["pos"] = centres[rng.integers(0, k, n)] + rng.normal(0, 0.15 * half, (n, 3))
    s = np.exp(rng.normal(np.log(0.01 * half), 0.6, (n, 3)))
    stretch = rng.random(n) < 0.3
    s[stretch, rng.integers(0, 3, int(stretch.sum()))] *= rng.uniform(1.0, 30.0, int(stretch.sum()))
    g["scale"] = np.minimum(s, 0.5 * half)
    q = rng.normal(size=(n, 4))
    g["rot"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    g["opacity"] = 1.0 / (1.0 + np.exp(-rng.normal(0, 2.0, n)))
    g["k_scale"] = 1.0
    g["f_dc"] = rng.normal(0, 1.0, (n, 3))
    if rng.random() < 0.5:
        g["f_rest"] = rng.normal(0, 0.3, (n, 45))
    return S.HostScene(g, K.Tree.build(g))


def _view(rng, aabb):
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), hi - lo
    if rng.random() < 0.3:
        eye = lo + ext * rng.uniform(0.2, 0.8, 3)  # inside the box
    else:
        d = rng.normal(size=3)
        eye = c + d / np.linalg.norm(d) * np.linalg.norm(ext) * rng.uniform(0.6, 2.0)
    target = c + ext * rng.uniform(-0.2, 0.2, 3)
    w, h = int(rng.integers(48, 161)), int(rng.integers(48, 161))
    fov = float(rng.uniform(20.0, 100.0))
    return K.camera_make(K.look_at(eye, target, fov, w / h, 0.005), w, h), w, h


def _rays(rng, aabb, n=3000):
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    ext = hi - lo
    o = (lo - 0.25 * ext + 1.5 * ext * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_view_and_rays(dev, stream, seed):
    rng = np.random.default_rng(7000 + seed)
    hs = _scene(rng)
    ci, w, h = _view(rng, hs.aabb)
    depth = int(rng.integers(2, 25))
    print(f"[sweep {seed}] {len(hs.g)} Gaussians, {w}x{h}, depth {depth}", flush=True)
    sc = dev.scene(hs.tree, hs.g)
    gpu = dev.render_debug(sc, ci, w, h, depth=depth, seq_cap=16)
    ref = O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h, depth=depth, seq_cap=16)
    assert np.array_equal(gpu["counters"], ref["counters"]), f"seed {seed}: counters differ"
    assert np.array_equal(gpu["hash"], ref["hash"]), f"seed {seed}: composited-ID hash differs"
    assert np.array_equal(gpu["seq"], ref["seq"]), f"seed {seed}: composited-ID sequence differs"
    assert np.array_equal(gpu["rgbt"].view(np.uint32), ref["rgbt"].reshape(h, w, 4).view(np.uint32)), \
        f"seed {seed}: RGBT differs"
    img = dev.render_image(sc, ci, w, h, depth=depth)
    assert np.array_equal(img.view(np.uint32), gpu["rgbt"].view(np.uint32)), f"seed {seed}: plain image differs"
    o, d = _rays(rng, hs.aabb)
    rr = dev.render_rays(sc, o, d, depth=depth)
    rd = dev.render_rays(sc, o, d, depth=depth, debug=True)
    rref = O.render_rays(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, o, d, depth=depth)
    assert np.array_equal(rd["counters"], rref["counters"]), f"seed {seed}: ray counters differ"
    assert np.array_equal(rd["hash"], rref["hash"]), f"seed {seed}: ray hashes differ"
    for r in (rr, rd):
        assert np.array_equal(r["rgbt"].view(np.uint32), rref["rgbt"].reshape(-1, 4).view(np.uint32)), \
            f"seed {seed}: ray RGBT differs"
    # some work happened: blends in the view or along the rays
    assert ref["counters"][..., 4].sum() + rref["counters"][..., 4].sum() > 0
    # the same view and rays as one persistent batch, random tuning
    tw, th = [(8, 8), (16, 8), (16, 16)][int(rng.integers(0, 3))]
    threads = int(rng.choice([64, 128])) if depth > 16 else int(rng.choice([64, 128, 256]))
    tun = K.Tuning(tw, th, depth, int(rng.choice([4, 7, 8, 16])), int(rng.choice([8, 16, 32, 48])), threads)
    outv, outr = dev.alloc(w * h * 16), dev.alloc(len(o) * 16)
    ob, db = dev.upload(o), dev.upload(d)
    b = K.Batch(dev, [K.View(sc.h, ci, w, h, outv.ptr)], [K.RaySet(sc.h, ob.ptr, db.ptr, len(o), outr.ptr)], tun)
    b.run(stream, use_graph=bool(rng.random() < 0.5))
    dev.sync(stream)
    bv, br = outv.download(np.float32, (h, w, 4)), outr.download(np.float32, (len(o), 4))
    b.free()
    for x in (outv, outr, ob, db):
        x.free()
    label = f"seed {seed}, tuning {tun.as_dict()}"
    assert np.array_equal(bv.view(np.uint32), ref["rgbt"].view(np.uint32)), f"{label}: batched view differs"
    assert np.array_equal(br.view(np.uint32), rref["rgbt"].view(np.uint32)), f"{label}: batched rays differ"
    sc.free()
