"""Serving with continuous batching into one running persistent kernel
(kdgrt_server_config.persistent_kernel = 1).

Each request is published into the queue of a k_serve kernel that is already
running; the kernel's waves take its tiles after those of earlier requests.
Every served RGBT image must equal the plain one-view launch of the same
request bit for bit (and one the FP32 oracle).  The cases cover requests
joining a running kernel, epochs that fill their request table and drain
into a new epoch, epochs stopped when idle and restarted by the next
request, and kernels whose waves give up waiting early (a tiny poll limit
set from the host), whose unfinished requests are published again.
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

RES = [(64, 48), (128, 96), (160, 120), (96, 96), (200, 150), (320, 240)]


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def three_scenes(dev):
    hss = [S.uniform_10k(), S.clustered(30000, seed=1003), S.clustered(60000, seed=1004)]
    scs = [dev.scene(h.tree, h.g) for h in hss]
    yield hss, scs
    for sc in scs:
        sc.free()


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


def _serve(dev, scs, reqs, gaps=None, **kw):
    budget = sum(sc.device_bytes() for sc in scs) + (256 << 20)
    srv = K.Server(dev, budget, persistent=True, **kw)
    ids = [srv.add_resident_scene(sc) for sc in scs]
    srv.start()
    results = [None] * len(reqs)
    rids = [None] * len(reqs)

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
        if gaps is not None and gaps[i] > 0:
            time.sleep(gaps[i])
        rids[i] = srv.submit(ids[k], ci, w, h)
    th.join()
    st = srv.stats()
    srv.close()
    return results, st


def _check(dev, scs, reqs, results, label):
    for i, ((img, info), (k, ci, w, h)) in enumerate(zip(results, reqs)):
        assert info.status == 0, f"{label}: request {i} failed with {info.status}"
        ref = dev.render_image(scs[k], ci, w, h)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), f"{label}: request {i} served != alone"


def test_persistent_requests_join_the_running_kernel(dev, three_scenes):
    hss, scs = three_scenes
    big = (2, S.cam_orbit(hss[2].aabb, 1280, 720, radius=25.0), 1280, 720)
    reqs = [big] + _requests(hss, 47, 21)
    gaps = [0.0] * len(reqs)
    gaps[24] = 0.003  # a second wave while the first runs
    print("[persistent] 48 requests, 3 scenes", flush=True)
    results, st = _serve(dev, scs, reqs, gaps=gaps)
    _check(dev, scs, reqs, results, "join")
    infos = [i for _, i in results]
    print(f"[persistent] epochs {st['batches']}, max requests in the kernel {max(i.batch_views for i in infos)}")
    assert st["completed"] == len(reqs)
    assert max(i.batch_views for i in infos) > 1, "no request joined a kernel running another"
    assert st["batches"] <= 3
    k, ci, w, h = reqs[0]
    ref = O.render(hss[k].nodes, hss[k].offsets, hss[k].g_inv, hss[k].aabb, ci, w, h)
    assert np.array_equal(results[0][0], ref["rgbt"])


def test_persistent_epochs_fill_and_drain(dev, three_scenes):
    hss, scs = three_scenes
    reqs = _requests(hss, 40, 22)
    print("[persistent] 40 requests, 8 per epoch", flush=True)
    results, st = _serve(dev, scs, reqs, epoch_max_requests=8)
    _check(dev, scs, reqs, results, "fill")
    assert st["batches"] >= 5, st


def test_persistent_idle_stop_and_restart(dev, three_scenes):
    hss, scs = three_scenes
    reqs = _requests(hss, 12, 23)
    gaps = [0.0] * 12
    gaps[4] = gaps[8] = 0.05   # longer than the idle time: the epoch stops in between
    print("[persistent] 12 requests in three bursts", flush=True)
    results, st = _serve(dev, scs, reqs, gaps=gaps, epoch_idle_ms=2.0)
    _check(dev, scs, reqs, results, "idle")
    assert st["batches"] >= 3, st


def test_persistent_waves_that_give_up_early(dev, three_scenes):
    """A poll limit of 200: waves leave the kernel after a short wait, so
    epochs end on their own; requests they left unfinished are published
    again and still come back equal to the plain launch."""
    hss, scs = three_scenes
    reqs = _requests(hss, 24, 24)
    gaps = [0.002 if i % 3 == 0 else 0.0 for i in range(24)]
    print("[persistent] 24 requests, poll limit 200", flush=True)
    results, st = _serve(dev, scs, reqs, gaps=gaps, spin_limit=200)
    _check(dev, scs, reqs, results, "early")
    assert st["completed"] == len(reqs)


def test_persistent_idle_plain(dev, three_scenes):
    """idle_plain: a request that finds the device idle runs as one plain
    one-view launch (batch_views 1); requests that arrive while it runs wait
    for it and then start an epoch of the persistent kernel together; lone
    requests after an idle stop run plain again.  Every result equals the
    plain launch of the same request bit for bit."""
    hss, scs = three_scenes
    big = (2, S.cam_orbit(hss[2].aabb, 1920, 1080, radius=25.0), 1920, 1080)
    reqs = [big] + _requests(hss, 23, 25) + [big]
    gaps = [0.0] * len(reqs)
    gaps[0] = 0.02
    gaps[1] = 0.001   # the scheduler takes the first request alone before the burst arrives
    gaps[-1] = 0.05   # after the idle stop: a lone request again
    print("[persistent] idle-plain launches and an epoch", flush=True)
    results, st = _serve(dev, scs, reqs, gaps=gaps, epoch_idle_ms=2.0, idle_plain=True)
    _check(dev, scs, reqs, results, "idle-plain")
    views = [i.batch_views for _, i in results]
    print(f"[persistent] requests per launch or epoch: {views}, launches + epochs {st['batches']}")
    assert st["completed"] == len(reqs)
    assert views[0] == 1 and views[-1] == 1, "a request on an idle device did not run alone"
    assert max(views) > 1, "the burst did not start an epoch of the persistent kernel"
    assert st["batches"] >= 3, st
