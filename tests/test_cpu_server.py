"""Serving loop on the modelled device (no GPU): admission by memory budget,
continuous batching, arrival order, deadlines, cancellation, accounting."""
import ctypes as C
import importlib
import threading
import time

import numpy as np
import pytest

K = importlib.import_module("3dgrt_kdtree_amd")


def small_tree():
    g = K.generate_gaussians(0, 2000, 5)
    return g, K.Tree.build(g)


def make_server(budget, rate_mrays, max_views=8, inflight=2, window_ms=2.0, plain=False):
    cfg = K.ServerConfig()
    cfg.hbm_budget_bytes = budget
    cfg.max_batch_views = max_views
    cfg.max_inflight_batches = inflight
    cfg.plain_launches = 1 if plain else 0
    cfg.tuning = K.default_tuning()
    cfg.batch_window_ms = window_ms
    cfg.simulate_mrays_per_s = rate_mrays
    s = C.c_void_p()
    K.check(K.lib().kdgrt_server_create(None, C.byref(cfg), C.byref(s)))
    return s


def add_scene(s, g, tree):
    sid = C.c_int32()
    K.check(K.lib().kdgrt_server_add_scene(s, tree.h, K._p(g), C.c_uint64(len(g)), 0, C.byref(sid)))
    return sid.value


def req(sid, w, h, deadline=0.0):
    r = K.Request()
    r.scene_id, r.width, r.height, r.arrival_s, r.deadline_s = sid, w, h, 0.0, deadline
    return r


def submit(s, r):
    rid = C.c_uint64()
    st = K.lib().kdgrt_server_submit(s, C.byref(r), C.byref(rid))
    return st, rid.value


def wait(s, rid, timeout_ms=5000):
    info = K.ResultInfo()
    st = K.lib().kdgrt_server_wait(s, C.c_uint64(rid), C.c_double(timeout_ms), None, C.byref(info))
    return st, info


def stats(s):
    st = K.ServerStats()
    K.check(K.lib().kdgrt_server_stats_get(s, C.byref(st)))
    return st


def test_batches_form_under_load_and_order_is_kept():
    g, tree = small_tree()
    s = make_server(1 << 30, rate_mrays=50.0, max_views=8, inflight=2)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    ids = []
    for i in range(40):
        st, rid = submit(s, req(sid, 320, 240))
        assert st == 0
        ids.append(rid)
    res = [wait(s, rid) for rid in ids]
    assert all(st == 0 for st, _ in res)
    infos = [i for _, i in res]
    assert max(i.batch_views for i in infos) > 1          # batching happened
    launched = [i.launched_s for i in infos]
    assert launched == sorted(launched)                     # admission in arrival order
    assert all(i.done_s >= i.launched_s >= i.arrival_s for i in infos)
    sst = stats(s)
    assert sst.completed == 40 and sst.rays == 40 * 320 * 240 and sst.batches < 40
    K.check(K.lib().kdgrt_server_destroy(s))


def test_budget_admission_and_rejection():
    g, tree = small_tree()
    scene_bytes = K.lib().kdgrt_scene_bytes_estimate(len(tree.nodes), len(tree.offsets), len(g))
    out_bytes = 640 * 480 * 16
    s = make_server(scene_bytes + 2 * out_bytes + 1024, rate_mrays=5.0, max_views=8, inflight=4)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    st, _ = submit(s, req(sid, 1920, 1080))
    assert st == 6  # KDGRT_ERR_CAPACITY: can never fit the arena
    ids = [submit(s, req(sid, 640, 480))[1] for _ in range(5)]
    time.sleep(0.05)
    sst = stats(s)
    assert sst.bytes_peak <= scene_bytes + 2 * out_bytes + 1024   # never more than two outputs admitted
    res = [wait(s, rid) for rid in ids]
    assert all(st == 0 for st, _ in res)
    infos = [i for _, i in res]
    assert max(i.batch_views for i in infos) <= 2
    K.check(K.lib().kdgrt_server_destroy(s))


def test_scene_over_budget_rejected():
    g, tree = small_tree()
    s = make_server(1000, rate_mrays=1.0)
    sid = C.c_int32()
    assert K.lib().kdgrt_server_add_scene(s, tree.h, K._p(g), C.c_uint64(len(g)), 0, C.byref(sid)) == 6
    K.check(K.lib().kdgrt_server_destroy(s))


def test_deadline_and_cancel():
    g, tree = small_tree()
    s = make_server(1 << 30, rate_mrays=0.5, max_views=1, inflight=1)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    _, a = submit(s, req(sid, 640, 480))          # occupies the modelled device ~0.6 s
    _, b = submit(s, req(sid, 64, 64, deadline=1e-3))  # expires while waiting
    _, c = submit(s, req(sid, 64, 64))
    assert K.lib().kdgrt_server_cancel(s, C.c_uint64(c)) == 0
    assert wait(s, b)[0] == 9 and wait(s, c)[0] == 9
    assert wait(s, a)[0] == 0
    sst = stats(s)
    assert sst.cancelled == 2 and sst.completed == 1
    K.check(K.lib().kdgrt_server_destroy(s))


def test_wait_timeout_keeps_request():
    g, tree = small_tree()
    s = make_server(1 << 30, rate_mrays=0.2, max_views=1, inflight=1)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    _, a = submit(s, req(sid, 320, 240))
    assert wait(s, a, timeout_ms=1.0)[0] == 10      # TIMEOUT, request stays
    assert wait(s, a, timeout_ms=-1)[0] == 0
    K.check(K.lib().kdgrt_server_destroy(s))


def test_concurrent_submitters():
    g, tree = small_tree()
    s = make_server(1 << 30, rate_mrays=200.0, max_views=16, inflight=3)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    out = []
    lock = threading.Lock()

    def client(n):
        for _ in range(n):
            st, rid = submit(s, req(sid, 160, 120))
            assert st == 0
            st, info = wait(s, rid)
            with lock:
                out.append((st, info.done_s - info.arrival_s))

    th = [threading.Thread(target=client, args=(25,)) for _ in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert len(out) == 100 and all(st == 0 for st, _ in out)
    assert stats(s).completed == 100
    K.check(K.lib().kdgrt_server_destroy(s))


def test_plain_mode_one_at_a_time():
    g, tree = small_tree()
    s = make_server(1 << 30, rate_mrays=20.0, max_views=8, inflight=4, plain=True)
    sid = add_scene(s, g, tree)
    K.check(K.lib().kdgrt_server_start(s))
    ids = [submit(s, req(sid, 320, 240))[1] for _ in range(10)]
    res = [wait(s, rid) for rid in ids]
    assert all(st == 0 for st, _ in res)
    infos = [i for _, i in res]
    assert all(i.batch_views == 1 for i in infos)
    for a, b in zip(infos, infos[1:]):
        assert b.launched_s >= a.done_s - 1e-9   # strictly one launch at a time
    K.check(K.lib().kdgrt_server_destroy(s))
