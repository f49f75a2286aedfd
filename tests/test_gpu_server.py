"""Serving loop on the GPU: every served result equals the same request
rendered alone.

Requests for three resident scenes arrive in two waves, the second while the
first is still on the device, so they are admitted into several overlapping
persistent batches (continuous batching).  Each served RGBT image must equal
the plain one-view launch of the same request bit for bit (the tolerance of
the path is 1e-5; the batched and plain kernels share the per-ray code, so
the observed difference is zero).
"""
import importlib
import threading
import time

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

RES = [(64, 48), (128, 96), (160, 120), (96, 96), (200, 150)]


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _requests(hss, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(0, len(hss)))
        w, h = RES[int(rng.integers(0, len(RES)))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        radius = 3.0 if k == 0 else 25.0
        out.append((k, S.cam_orbit(hss[k].aabb, w, h, radius=radius, direction=tuple(v)), w, h))
    return out


def _serve(dev, hss, reqs, plain, max_views=6, inflight=3, resident=None):
    """resident: device Scenes to serve instead of uploading the host trees."""
    if resident is None:
        budget = sum(K.lib().kdgrt_scene_bytes_estimate(len(h.nodes), len(h.offsets), len(h.g)) for h in hss)
    else:
        budget = sum(sc.device_bytes() for sc in resident)
    srv = K.Server(dev, budget + (64 << 20), max_batch_views=max_views, max_inflight=inflight, plain=plain,
                   window_ms=1.0)
    if resident is None:
        ids = [srv.add_scene(h.tree, h.g) for h in hss]
    else:
        ids = [srv.add_resident_scene(sc) for sc in resident]
    srv.start()
    results = [None] * len(reqs)
    rids = [None] * len(reqs)
    half = len(reqs) // 2

    def waiter():
        for i in range(len(reqs)):
            while rids[i] is None:
                time.sleep(1e-4)
            k, ci, w, h = reqs[i]
            img = np.zeros((h, w, 4), np.float32)
            info = srv.wait(rids[i], timeout_ms=60000, out=img)
            results[i] = (img, info)

    th = threading.Thread(target=waiter)
    th.start()
    for i, (k, ci, w, h) in enumerate(reqs):
        if i == half:
            time.sleep(0.003)  # second wave arrives while the first runs
        rids[i] = srv.submit(ids[k], ci, w, h)
    th.join()
    st = srv.stats()
    srv.close()
    return results, st


def _alone(dev, hss, reqs):
    scs = [dev.scene(h.tree, h.g) for h in hss]
    out = [dev.render_image(scs[k], ci, w, h) for k, ci, w, h in reqs]
    for sc in scs:
        sc.free()
    return out


@pytest.fixture(scope="module")
def three_scenes():
    return [S.uniform_10k(), S.clustered(30000, seed=1003), S.clustered(60000, seed=1004)]


def test_served_equals_alone(dev, three_scenes):
    hss = three_scenes
    # a large view first: the device is busy with it while the rest arrive
    big = (2, S.cam_orbit(hss[2].aabb, 1280, 720, radius=25.0), 1280, 720)
    reqs = [big] + _requests(hss, 47, 11)
    print("[server] 48 requests, 3 scenes, batched", flush=True)
    results, st = _serve(dev, hss, reqs, plain=False)
    alone = _alone(dev, hss, reqs)
    for i, ((img, info), ref) in enumerate(zip(results, alone)):
        assert info.status == 0, f"request {i} failed with {info.status}"
        assert np.array_equal(img, ref), f"request {i}: served != alone, max {np.abs(img - ref).max()}"
    assert st["completed"] == len(reqs) and st["rejected"] == 0
    infos = [info for _, info in results]
    assert max(i.batch_views for i in infos) > 1, "no request was batched with another"
    # continuous batching: some batch was launched while an earlier one was still running
    by_batch = {}
    for i in infos:
        by_batch.setdefault(i.batch_id, (i.launched_s, i.done_s))
    spans = sorted(by_batch.values())
    overlap = any(spans[j][0] < spans[j - 1][1] for j in range(1, len(spans)))
    print(f"[server] batches {len(spans)}, max views {max(i.batch_views for i in infos)}, overlap {overlap}")
    assert overlap, "no two batches overlapped on the device"
    # one served image against the oracle as well
    k, ci, w, h = reqs[0]
    ref = O.render(hss[k].nodes, hss[k].offsets, hss[k].g_inv, hss[k].aabb, ci, w, h)
    assert np.abs(results[0][0].astype(np.float64) - ref["rgbt"]).max() <= 1e-5


def test_served_plain_mode_equals_alone(dev, three_scenes):
    hss = three_scenes
    reqs = _requests(hss, 12, 12)
    print("[server] 12 requests, plain mode", flush=True)
    results, st = _serve(dev, hss, reqs, plain=True)
    alone = _alone(dev, hss, reqs)
    for i, ((img, info), ref) in enumerate(zip(results, alone)):
        assert info.status == 0
        assert info.batch_views == 1
        assert np.array_equal(img, ref), f"request {i}: served != alone"
    assert st["completed"] == len(reqs)


def test_served_resident_scenes_equal_alone(dev, three_scenes):
    """Scenes the caller keeps resident (kdgrt_server_add_resident_scene)
    serve the same images; the server does not free them."""
    hss = three_scenes
    scs = [dev.scene(h.tree, h.g) for h in hss]
    reqs = _requests(hss, 16, 13)
    print("[server] 16 requests, resident scenes", flush=True)
    results, st = _serve(dev, hss, reqs, plain=False, resident=scs)
    for i, ((img, info), (k, ci, w, h)) in enumerate(zip(results, reqs)):
        assert info.status == 0
        assert np.array_equal(img, dev.render_image(scs[k], ci, w, h)), f"request {i}: served != alone"
    assert st["completed"] == len(reqs)
    for sc in scs:
        sc.free()
