"""Correctness at the full measured workloads: the 64-view batch on the 1M
scene, the 8-scene batch with 4M incoherent rays, and Poisson serving over
16 resident scenes.  The same generators as the benchmark build them.

Bar (per ray, cudaRenderer.cu:2313-2390): RGB and T bit-identical to the FP32
oracle for every compared ray; served results bit-identical to the same
request rendered alone.
  * 64 views, 1M Gaussians (48M rays): every view of the autotuned persistent
    launch, replayed as a hipGraph, against the oracle.
  * 8 scenes (Gaussian counts log-uniform in [100k, 3M]), 256 views + 4M
    incoherent rays in one launch: 32 views and 200,000 of the rays against
    the oracle, every output of the batch against the plain launches.
  * 16 resident scenes, Poisson arrivals at 1000 requests/s for one second:
    every served image against the plain launch of the same request, and 24
    of them against the oracle.
"""
import hashlib
import importlib
import threading
import time

import numpy as np
import pytest

import bench as B
import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def test_c2_64_views_batched(dev):
    hs = S.clustered(1_000_000)
    sc = dev.scene(hs.tree, hs.g)
    views = B.views_64(hs.aabb)
    bufs, vlist = [], []
    for ci, w, h in views:
        b = dev.alloc(w * h * 16)
        bufs.append(b)
        vlist.append(K.View(sc.h, ci, w, h, b.ptr))
    print(f"[c2] {sum(w * h for _, w, h in views)} rays; autotune", flush=True)
    tun, _ = K.autotune(dev, vlist, max_trials=8, reps=1)
    batch = K.Batch(dev, vlist, (), tun)
    stream = dev.stream()
    nan = {(w, h): np.full((h, w, 4), np.nan, np.float32) for _, w, h in views}
    for rep in range(2):
        for (ci, w, h), b in zip(views, bufs):
            b.upload(nan[(w, h)])
        batch.run(stream, use_graph=True)
        dev.sync(stream)
    print(f"[c2] tuning {tun.as_dict()}; oracle", flush=True)
    for k, ((ci, w, h), b) in enumerate(zip(views, bufs)):
        got = b.download(np.float32, (h, w, 4))
        ref = O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h)["rgbt"]
        assert _bits_equal(got, ref), f"c2 view {k} ({w}x{h}): max |d| {np.nanmax(np.abs(got - ref))}"
    batch.free()
    for b in bufs:
        b.free()
    sc.free()


@pytest.fixture(scope="module")
def scene_set16(dev):
    """The 16 scenes of the multi-scene and serving workloads, with their host
    trees kept for the oracle."""
    rng = np.random.default_rng(1004)
    counts = np.exp(rng.uniform(np.log(1e5), np.log(3e6), 16)).astype(np.int64)
    out = []
    t0 = time.time()
    for j, n in enumerate(counts):
        g = K.generate_gaussians(1, int(n), 1004 * 1000 + j)
        tree = K.Tree.build(g)
        out.append((S.HostScene(g, tree), dev.scene(tree, g)))
        print(f"[scenes] {j + 1}/16: {n} Gaussians ({time.time() - t0:.0f} s)", flush=True)
    yield out
    for _, sc in out:
        sc.free()


def test_c3_multi_scene_batch(dev, scene_set16):
    scenes = scene_set16[:8]
    rng = np.random.default_rng(1003)
    views, vlist, bufs = [], [], []
    for _ in range(256):
        k = int(rng.integers(0, len(scenes)))
        ci, w, h = B.random_view(rng, scenes[k][0].aabb)
        b = dev.alloc(w * h * 16)
        bufs.append(b)
        views.append((k, ci, w, h, b))
        vlist.append(K.View(scenes[k][1].h, ci, w, h, b.ptr))
    per = 4_000_000 // len(scenes)
    sets, rays = [], []
    for k, (hs, sc) in enumerate(scenes):
        lo, hi = hs.aabb[:3].astype(np.float64), hs.aabb[3:].astype(np.float64)
        o = (lo + (hi - lo) * rng.random((per, 3))).astype(np.float32)
        d = rng.normal(size=(per, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        ob, db, out = dev.upload(o), dev.upload(d), dev.alloc(per * 16)
        bufs += [ob, db, out]
        sets.append(K.RaySet(sc.h, ob.ptr, db.ptr, per, out.ptr))
        rays.append((o, d, out))
    print(f"[c3] 256 views + {per * len(scenes)} rays: batch", flush=True)
    batch = K.Batch(dev, vlist, sets, K.default_tuning(dev))
    stream = dev.stream()
    batch.run(stream, use_graph=True)
    batch.run(stream, use_graph=True)
    dev.sync(stream)
    # every output against the plain launches of the same views and ray lists
    for k, ci, w, h, b in views:
        got = b.download(np.float32, (h, w, 4))
        assert _bits_equal(got, dev.render_image(scenes[k][1], ci, w, h)), f"c3 view of scene {k} {w}x{h}"
    for k, (o, d, out) in enumerate(rays):
        got = out.download(np.float32, (per, 4))
        assert _bits_equal(got, dev.render_rays(scenes[k][1], o, d)["rgbt"]), f"c3 ray set {k}"
    print("[c3] oracle sample", flush=True)
    for k, ci, w, h, b in views[:32]:
        hs = scenes[k][0]
        ref = O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h)["rgbt"]
        assert _bits_equal(b.download(np.float32, (h, w, 4)), ref), f"c3 view of scene {k} vs oracle"
    for k, (o, d, out) in enumerate(rays):
        hs = scenes[k][0]
        idx = np.random.default_rng(k).choice(per, 25_000, replace=False)
        ref = O.render_rays(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, o[idx], d[idx])["rgbt"]
        assert _bits_equal(out.download(np.float32, (per, 4))[idx], ref), f"c3 ray set {k} vs oracle"
    batch.free()
    for b in bufs:
        b.free()


def test_c4_serving_poisson_1000(dev, scene_set16):
    scenes = scene_set16
    resident = sum(sc.device_bytes() for _, sc in scenes)
    srv = K.Server(dev, int(resident / 0.8), max_batch_views=32, max_inflight=3, plain=False, window_ms=0.0)
    ids = [srv.add_resident_scene(sc) for _, sc in scenes]
    srv.start()
    rng = np.random.default_rng(3000)
    arrivals = np.cumsum(rng.exponential(1e-3, 1500))
    arrivals = arrivals[arrivals < 1.0]
    reqs = []
    for _ in arrivals:
        k = int(rng.integers(0, len(scenes)))
        ci, w, h = B.random_view(rng, scenes[k][0].aabb)
        reqs.append((k, ci, w, h))
    rids = [None] * len(reqs)
    digests = [None] * len(reqs)
    views = [0] * len(reqs)
    keep = {}

    def waiter():
        for i, (k, ci, w, h) in enumerate(reqs):
            while rids[i] is None:
                time.sleep(2e-4)
            img = np.empty((h, w, 4), np.float32)
            info = srv.wait(rids[i], timeout_ms=-1.0, out=img)
            assert info.status == 0
            views[i] = info.batch_views
            digests[i] = hashlib.blake2b(img.tobytes(), digest_size=16).digest()
            if i < 24:
                keep[i] = img

    th = threading.Thread(target=waiter)
    th.start()
    print(f"[c4] {len(reqs)} requests over 1 s", flush=True)
    t0 = time.perf_counter()
    for i, (k, ci, w, h) in enumerate(reqs):
        while time.perf_counter() < t0 + arrivals[i]:
            time.sleep(1e-4)
        rids[i] = srv.submit(ids[k], ci, w, h)
    th.join()
    srv.close()
    assert all(dg is not None for dg in digests)
    assert max(views) > 1, "no request was served in a batch of several views"
    print("[c4] alone renders", flush=True)
    for i, (k, ci, w, h) in enumerate(reqs):
        alone = dev.render_image(scenes[k][1], ci, w, h)
        assert hashlib.blake2b(alone.tobytes(), digest_size=16).digest() == digests[i], f"c4 request {i}"
    print("[c4] oracle sample", flush=True)
    for i, img in keep.items():
        k, ci, w, h = reqs[i]
        hs = scenes[k][0]
        assert _bits_equal(img, O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, ci, w, h)["rgbt"]), f"c4 {i}"
