"""Benchmark of the kd-tree 3D-Gaussian ray-render hot path on one MI355X.

Headline line (one JSON object on stdout): the plain one-view launch over the
1M-Gaussian clustered scene at 1920x1080 (1 ray/pixel), Mrays/s with the
scene resident in HBM, plus
  roofline      compulsory bytes of a render call (every record its rays read,
                counted once, from the device touch maps) / mean call time vs
                8 TB/s; the per-ray logical bytes of the reference layout and
                the measured HBM traffic beside it,
  valu          the VALU-issue rate of k_view against the measured ceiling of
                a pure fused-multiply-add kernel,
  texture       k_view's texture-path busy share and the share that its load
                instructions fill at the path's 16-cycle floor per instruction
                (counter runs summarised under profiles/),
  cpu_baseline  the FP32 oracle (CPU restatement of the reference
                traversal+compositing, OpenMP) on this host's cores, and the
                10k-Gaussian 256x256 case the reference runs on the CPU,
  batched       the 64 mixed-resolution views in one persistent launch
                (hipGraph replay) against 64 plain launches of the same views,
  hbm_copy      a 1 GiB streaming copy (k_copy), the measured HBM ceiling and
                the calibration kernel of the traffic counters,
  build         host SAH build time of the scene.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

K = importlib.import_module("3dgrt_kdtree_amd")

METRIC = "Mrays/s + HBM roofline % @1M Gaussians; p50/p99 view latency, served vs plain"
HBM_PEAK_GBS = 8000.0
RESOLUTIONS = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def scene_1m(n=1_000_000, seed=1001):
    g = K.generate_gaussians(1, n, seed)
    t0 = time.time()
    tree = K.Tree.build(g)
    return g, tree, time.time() - t0


def view_1080p(aabb):
    centre = 0.5 * (aabb[:3].astype(np.float64) + aabb[3:].astype(np.float64))
    d = np.array([0.48, 0.36, 0.8])
    d /= np.linalg.norm(d)
    cam = K.look_at(centre + 25.0 * d, centre, 60.0, 1920 / 1080, 0.005)
    return K.camera_make(cam, 1920, 1080)


def views_64(aabb, seed=1002):
    """64 views: resolution uniform over the four sizes, eye uniform on the
    sphere of radius 25 about the box centre, aim jittered by +-10 degrees."""
    rng = np.random.default_rng(seed)
    centre = 0.5 * (aabb[:3].astype(np.float64) + aabb[3:].astype(np.float64))
    out = []
    for _ in range(64):
        w, h = RESOLUTIONS[rng.integers(0, 4)]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = centre + 25.0 * d
        fwd = centre - eye
        fwd /= np.linalg.norm(fwd)
        yaw, pitch = np.radians(rng.uniform(-10, 10, 2))
        up = np.array([0.0, 1.0, 0.0]) if abs(fwd[1]) < 0.99 else np.array([0.0, 0.0, 1.0])
        right = np.cross(fwd, up)
        right /= np.linalg.norm(right)
        upv = np.cross(right, fwd)
        fwd2 = fwd * math.cos(yaw) * math.cos(pitch) + right * math.sin(yaw) + upv * math.sin(pitch)
        cam = K.look_at(eye, eye + fwd2, 60.0, w / h, 0.005)
        out.append((K.camera_make(cam, w, h), w, h))
    return out


def algorithmic_bytes(counters):
    """B_ray = 8*(node+leaf visits) + 52*tests + 256*evaluated hits + 16 (RGBA out),
    reference data layout, summed over rays."""
    c = counters.reshape(-1, K.N_COUNTERS).astype(np.int64)
    return int(8 * (c[:, 0] + c[:, 1]).sum() + 52 * c[:, 2].sum() + 256 * c[:, 6].sum() + 16 * len(c))


def load_profile(name, workload_key):
    """One entry of a counter summary under profiles/ (None when absent)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(workload_key)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_bandwidth(dev, stream, gib=1, reps=5):
    """Streaming copy of `gib` GiB (k_copy): read + write bytes / event time."""
    import ctypes as C
    n4 = gib * (1 << 30) // 16
    src, dst = dev.alloc(n4 * 16), dev.alloc(n4 * 16)
    K.check(K.lib().kdgrt_memset(dev.h, src.ptr, 0, C.c_size_t(n4 * 16)))
    K.check(K.lib().kdgrt_copy_f32x4(dev.h, dst.ptr, src.ptr, C.c_uint64(n4), stream))
    dev.sync()
    dev.timer_begin(stream)
    for _ in range(reps):
        K.check(K.lib().kdgrt_copy_f32x4(dev.h, dst.ptr, src.ptr, C.c_uint64(n4), stream))
    ms = dev.timer_end(stream) / reps
    src.free()
    dst.free()
    return {"kernel": "k_copy", "bytes_read": n4 * 16, "bytes_written": n4 * 16, "ms": round(ms, 4),
            "GB_s": round(2 * n4 * 16 / (ms * 1e-3) / 1e9, 1)}


def random_view(rng, aabb):
    """A request of the batched-view mix: resolution uniform over the four
    sizes, eye uniform on the radius-25 sphere about the box centre, aim
    jittered by +-10 degrees, 60 degree FOV."""
    centre = 0.5 * (aabb[:3].astype(np.float64) + aabb[3:].astype(np.float64))
    w, h = RESOLUTIONS[rng.integers(0, 4)]
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    eye = centre + 25.0 * d
    fwd = centre - eye
    fwd /= np.linalg.norm(fwd)
    yaw, pitch = np.radians(rng.uniform(-10, 10, 2))
    up = np.array([0.0, 1.0, 0.0]) if abs(fwd[1]) < 0.99 else np.array([0.0, 0.0, 1.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    fwd2 = fwd * math.cos(yaw) * math.cos(pitch) + right * math.sin(yaw) + upv * math.sin(pitch)
    return K.camera_make(K.look_at(eye, eye + fwd2, 60.0, w / h, 0.005), w, h), w, h


def run_poisson(srv, ids, aabbs, lam, duration_s, seed):
    """Poisson arrivals at `lam` requests/s for `duration_s`; each request a
    random resident scene and a random view of the mix.  Latency is arrival
    to result ready (server clock)."""
    import threading
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / lam, size=int(lam * duration_s * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration_s]
    reqs = []
    for _ in range(len(arrivals)):
        k = int(rng.integers(0, len(ids)))
        ci, w, h = random_view(rng, aabbs[k])
        reqs.append((ids[k], ci, w, h))
    rids = [None] * len(reqs)
    infos = [None] * len(reqs)

    def waiter():
        for i in range(len(reqs)):
            while rids[i] is None:
                time.sleep(2e-4)
            infos[i] = srv.wait(rids[i], timeout_ms=-1.0)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i, (sid, ci, w, h) in enumerate(reqs):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 2e-3:
            time.sleep(delay - 1e-3)
        while time.perf_counter() < t0 + arrivals[i]:
            pass
        rids[i] = srv.submit(sid, ci, w, h)
    th.join()
    lat = np.array([(x.done_s - x.arrival_s) * 1e3 for x in infos])
    rays = sum(w * h for _, _, w, h in reqs)
    span = max(x.done_s for x in infos) - min(x.arrival_s for x in infos)
    views = np.array([x.batch_views for x in infos])
    return {"requests": len(reqs), "offered_mrays_s": round(rays / duration_s / 1e6, 1),
            "mrays_s": round(rays / span / 1e6, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3),
            "mean_views_per_launch": round(float(views.mean()), 2),
            "max_views_per_launch": int(views.max()),
            "failed": int(sum(1 for x in infos if x.status != 0))}


def scene_set(dev, n_scenes, seed=1004):
    """n_scenes device scenes from the clustered generator with Gaussian
    counts log-uniform in [100k, 3M] (the multi-scene and serving sets)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(1e5), np.log(3e6), n_scenes)).astype(np.int64)
    t0 = time.time()
    scenes, aabbs = [], []
    for j, n in enumerate(counts):
        g = K.generate_gaussians(1, int(n), seed * 1000 + j)
        tree = K.Tree.build(g)
        scenes.append(dev.scene(tree, g))
        aabbs.append(tree.aabb.copy())
        tree.free()
        del g
    return scenes, aabbs, counts, time.time() - t0


def multi_scene_batch(dev, scenes, aabbs, counts, tuning, n_views=256, n_rays=4_000_000, seed=1003, reps=3):
    """Multi-scene batch: n_views views of the resolution mix over the scenes
    plus n_rays incoherent rays (origins uniform in a scene's box, directions
    uniform on the sphere) split over the scenes, in one persistent launch,
    against one plain launch per view / ray list, each waited for."""
    rng = np.random.default_rng(seed)
    views, plain, bufs = [], [], []
    rays_total = 0
    for _ in range(n_views):
        k = int(rng.integers(0, len(scenes)))
        ci, w, h = random_view(rng, aabbs[k])
        b = dev.alloc(w * h * 16)
        bufs.append(b)
        views.append(K.View(scenes[k].h, ci, w, h, b.ptr))
        plain.append(("view", scenes[k], ci, w, h, b))
        rays_total += w * h
    sets = []
    per = n_rays // len(scenes)
    for k, sc in enumerate(scenes):
        lo, hi = aabbs[k][:3].astype(np.float64), aabbs[k][3:].astype(np.float64)
        o = (lo + (hi - lo) * rng.random((per, 3))).astype(np.float32)
        d = rng.normal(size=(per, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        ob, db, out = dev.upload(o), dev.upload(d), dev.alloc(per * 16)
        bufs += [ob, db, out]
        sets.append(K.RaySet(sc.h, ob.ptr, db.ptr, per, out.ptr))
        plain.append(("rays", sc, ob, db, per, out))
        rays_total += per
    stream = dev.stream()

    def run_plain(sync_each=True):
        for p in plain:
            if p[0] == "view":
                _, sc, ci, w, h, b = p
                dev.render(sc, ci, w, h, b, stream=stream)
            else:
                _, sc, ob, db, n, out = p
                K.check(K.lib().kdgrt_render_rays(dev.h, sc.h, ob.ptr, db.ptr, K.C.c_uint64(n), 12, out.ptr,
                                                  stream))
            if sync_each:
                dev.sync(stream)
        dev.sync(stream)

    run_plain()
    t0 = time.perf_counter()
    for _ in range(reps):
        run_plain()
    plain_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        run_plain(sync_each=False)
    plain_nosync_s = (time.perf_counter() - t0) / reps
    batch = K.Batch(dev, views, sets, tuning)
    batch.run(stream, use_graph=True)
    dev.sync(stream)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run(stream, use_graph=True)
        dev.sync(stream)
    batch_s = (time.perf_counter() - t0) / reps
    batch.free()
    for b in bufs:
        b.free()
    out = {"scenes": len(scenes), "gaussians": [int(c) for c in counts], "views": n_views,
           "incoherent_rays": per * len(scenes), "rays": rays_total,
           "plain_mrays_s": round(rays_total / plain_s / 1e6, 2),
           "plain_nosync_mrays_s": round(rays_total / plain_nosync_s / 1e6, 2),
           "batched_graph_mrays_s": round(rays_total / batch_s / 1e6, 2), "speedup": round(plain_s / batch_s, 3),
           "speedup_vs_nosync": round(plain_nosync_s / batch_s, 3),
           "note": "every batched run sorts its ray sets into ray-key order before the launch (timed)",
           "tuning": tuning.as_dict()}
    log(f"multi-scene batch: {out}")
    return out


def serve_bench(dev, scenes, aabbs, counts, tuning, rates=((200, 4.0), (1000, 2.0)), window_ms=0.0):
    """Serving: the resident scenes (HBM budget = resident bytes / 0.8),
    Poisson arrivals, one view per request, in three modes: continuous
    batching into one running persistent kernel ("persistent"; a request
    that finds the device idle runs as one plain launch), overlapping
    batched launches on three streams ("batched"), and plain one-launch-per-
    request one at a time ("plain", the reference's way)."""
    resident = sum(sc.device_bytes() for sc in scenes)
    budget = int(resident / 0.8)
    log(f"serving: {len(scenes)} scenes, {int(counts.sum())} Gaussians, resident {resident / 1e9:.2f} GB, "
        f"budget {budget / 1e9:.2f} GB")
    out = {"scenes": len(scenes), "gaussians": [int(c) for c in counts], "resident_bytes": resident,
           "hbm_budget_bytes": budget, "batch_window_ms": window_ms, "max_batch_views": 32, "inflight_batches": 3}
    for mode in ("persistent", "batched", "plain"):
        srv = K.Server(dev, budget, max_batch_views=32, max_inflight=3, plain=(mode == "plain"), tuning=tuning,
                       window_ms=window_ms, persistent=(mode == "persistent"), idle_plain=(mode == "persistent"))
        ids = [srv.add_resident_scene(sc) for sc in scenes]
        srv.start()
        for lam, dur in rates:
            r = run_poisson(srv, ids, aabbs, lam, dur, seed=2000 + lam)
            out.setdefault(f"poisson_{lam}", {})[mode] = r
            log(f"serving {mode} at {lam} req/s: {r}")
        srv.close()
    for lam, _ in rates:
        e = out[f"poisson_{lam}"]
        e["served_vs_plain_mrays"] = round(e["batched"]["mrays_s"] / e["plain"]["mrays_s"], 3)
        e["persistent_vs_plain_mrays"] = round(e["persistent"]["mrays_s"] / e["plain"]["mrays_s"], 3)
    return out


def hybrid_spheres(aabb):
    """Secondary spheres of the hybrid leg, sized to the scene box: glass
    (ior 1.3), tinted mirror (reflectivity 0.85) and diffuse, around the
    scene centre."""
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    c, ext = 0.5 * (lo + hi), float((hi - lo).max())
    a = np.zeros(3, K.SECONDARY_SPHERE_DTYPE)
    a[0] = (c + 0.10 * ext * np.array([1.0, 0.3, 0.2]), 0.08 * ext, K.MATERIAL_GLASS, 1.0, 1.3, (0.95, 0.9, 0.85))
    a[1] = (c - 0.12 * ext * np.array([1.0, -0.2, 0.1]), 0.10 * ext, K.MATERIAL_MIRROR, 0.85, 1.0, (0.6, 0.7, 0.9))
    a[2] = (c + 0.20 * ext * np.array([0.0, 1.0, 0.0]), 0.04 * ext, K.MATERIAL_DIFFUSE, 0.0, 1.0, (0.8, 0.2, 0.1))
    return a


def hybrid_bench(dev, sc, ci, aabb, W, H, stream, steps, warmup):
    """The hybrid view (k_hybrid) of the same scene and camera with three
    secondary spheres in view, timed like the plain view."""
    sph = hybrid_spheres(aabb)
    buf = dev.alloc(W * H * 12)
    for _ in range(warmup):
        dev.render_hybrid(sc, ci, W, H, spheres=sph, out=buf, stream=stream)
    dev.sync()
    dev.timer_begin(stream)
    for _ in range(steps):
        dev.render_hybrid(sc, ci, W, H, spheres=sph, out=buf, stream=stream)
    ms = dev.timer_end(stream) / steps
    dev.sync()
    buf.free()
    r = {"workload": "hybrid view: same scene and 1920x1080 view, 3 secondary spheres (glass, mirror, diffuse) "
                     "around the scene centre, up to 5 segments per pixel",
         "ms_per_launch": round(ms, 4), "primary_mrays_s": round(W * H / (ms * 1e-3) / 1e6, 2)}
    log(f"hybrid 1080p: {ms:.3f} ms/launch, {r['primary_mrays_s']} primary Mrays/s")
    return r


def viewer_bench(dev, sc, aabb, W, H, frames=300, window=(100, 300)):
    """The interactive viewer's frame loop (kdgrt_viewer_*) at 1920x1080: the
    camera orbits the scene centre about the vertical at the reference's
    0.8 rad/s with 60 frames/s steps (0.76 degrees per frame), each frame is
    rendered and copied to pinned host memory while the next renders.  FPS as
    the reference reports it (1000 / render ms, averaged over the window of
    frames) and the delivered frame rate including the copies."""
    centre = 0.5 * (aabb[:3].astype(np.float64) + aabb[3:].astype(np.float64))
    d = np.array([0.48, 0.36, 0.8])
    d /= np.linalg.norm(d)
    out = {"workload": f"viewer frame loop: {W}x{H}, orbit 0.8 rad/s at dt 1/60 s about the scene centre, "
                       f"{frames} frames, FPS window {list(window)}, 3 slots"}
    for name, fmt in (("rgba8", K.FRAME_RGBA8), ("rgb_f32", K.FRAME_RGB_F32)):
        cam = K.look_at(centre + 25.0 * d, centre, 60.0, W / H, 0.005)
        v = K.Viewer(dev, sc, W, H, fmt=fmt, slots=3, window=window)
        ahead = 0
        t0 = time.perf_counter()
        for _ in range(frames):
            K.camera_orbit(cam, 0.8 / 60.0, tuple(centre), (0.0, 1.0, 0.0))
            v.submit(cam)
            ahead += 1
            if ahead == 2:
                f, _ = v.acquire(copy=False)
                v.release(f.frame_id)
                ahead -= 1
        while ahead:
            f, _ = v.acquire(copy=False)
            v.release(f.frame_id)
            ahead -= 1
        wall = time.perf_counter() - t0
        st = v.stats()
        v.close()
        out[name] = {"frame_bytes": W * H * (4 if fmt == K.FRAME_RGBA8 else 12),
                     "window_avg_fps": round(st["window_avg_fps"], 1),
                     "mean_render_ms": round(st["mean_render_ms"], 4),
                     "delivered_fps": round(frames / wall, 1)}
    log(f"viewer: {out}")
    return out


def orbit_view(aabb, yaw_deg, w=1920, h=1080):
    """view_1080p's camera turned by yaw_deg about the vertical through the
    box centre."""
    centre = 0.5 * (aabb[:3].astype(np.float64) + aabb[3:].astype(np.float64))
    d = np.array([0.48, 0.36, 0.8])
    d /= np.linalg.norm(d)
    a = np.radians(yaw_deg)
    d = np.array([d[0] * np.cos(a) + d[2] * np.sin(a), d[1], -d[0] * np.sin(a) + d[2] * np.cos(a)])
    cam = K.look_at(centre + 25.0 * d, centre, 60.0, w / h, 0.005)
    return K.camera_make(cam, w, h)


def tile_order_legs(dev, sc, ci, aabb, W, H, out, stream, steps):
    """The plain view in raster tile order (schedule off), and a camera that
    orbits by 0.1, 0.25 and 2 degrees per launch with and without the
    schedule.  For a camera that stays put the order comes from the previous
    launch's tile durations; once it moves (more than 0.05 degrees per
    launch), from the density of Gaussian centres projected into the image."""
    def timed(cams, sched):
        sc.set_tile_schedule(sched)
        dev.render(sc, cams[0], W, H, out, stream=stream)
        dev.sync()
        dev.timer_begin(stream)
        for c in cams:
            dev.render(sc, c, W, H, out, stream=stream)
        ms = dev.timer_end(stream) / len(cams)
        dev.sync()
        return ms
    r = {}
    ms = timed([ci] * steps, False)
    r["raster_order"] = {"ms_per_launch": round(ms, 4), "mrays_s": round(W * H / (ms * 1e-3) / 1e6, 1)}
    for step in (0.1, 0.25, 2.0):
        cams = [orbit_view(aabb, step * k, W, H) for k in range(steps)]
        m_sched, m_raster = timed(cams, True), timed(cams, False)
        r[f"orbit_{step}_deg_per_launch"] = {"scheduled_mrays_s": round(W * H / (m_sched * 1e-3) / 1e6, 1),
                                             "raster_mrays_s": round(W * H / (m_raster * 1e-3) / 1e6, 1)}
    sc.set_tile_schedule(True)
    log(f"tile order: {r}")
    return r


def rays_leg(dev, sc, aabb, n, stream, steps, warmup):
    """kdgrt_render_rays over n rays with origins uniform in the scene box and
    directions uniform on the sphere (the ray-key sort included)."""
    rng = np.random.default_rng(1005)
    lo, hi = aabb[:3].astype(np.float64), aabb[3:].astype(np.float64)
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    ob, db, out = dev.upload(o), dev.upload(d), dev.alloc(n * 16)

    def launch():
        K.check(K.lib().kdgrt_render_rays(dev.h, sc.h, ob.ptr, db.ptr, K.C.c_uint64(n), 12, out.ptr, stream))
    for _ in range(warmup):
        launch()
    dev.sync()
    dev.timer_begin(stream)
    for _ in range(steps):
        launch()
    ms = dev.timer_end(stream) / steps
    dev.sync()
    for b in (ob, db, out):
        b.free()
    r = {"rays": n, "ms_per_call": round(ms, 4), "mrays_s": round(n / (ms * 1e-3) / 1e6, 2)}
    log(f"incoherent rays: {r}")
    return r


def cpu_baseline(tree, g_inv, ci, w, h, budget_s):
    """The FP32 oracle with OpenMP on this host: the headline view (median of
    up to five full renders within budget_s), and the 10k-Gaussian 256x256
    case (uniform cube, eye (0,0,3)) that the reference runs on the CPU."""
    import oracle_lib as O
    import scenes as S
    threads = O.lib().oracle_max_threads()
    nodes, offs, aabb = tree.nodes, tree.offsets, tree.aabb
    times = []
    t_start = time.time()
    while len(times) < 5 and (time.time() - t_start) < budget_s:
        t0 = time.time()
        O.render(nodes, offs, g_inv, aabb, ci, w, h, threads=threads)
        times.append(time.time() - t0)
    med = float(np.median(times))
    g0 = K.generate_gaussians(0, 10000, 1000)
    t0 = time.time()
    tree0 = K.Tree.build(g0)
    build0 = time.time() - t0
    gi0, ci0 = K.invert_scales(g0), S.cam_c0(256, 256)
    t0s = []
    for _ in range(5):
        t0 = time.time()
        O.render(tree0.nodes, tree0.offsets, gi0, tree0.aabb, ci0, 256, 256, threads=threads)
        t0s.append(time.time() - t0)
    med0 = float(np.median(t0s))
    tree0.free()
    return {"value": round(w * h / med / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"full {w}x{h} view of the same 1M scene, FP32 oracle with OpenMP dynamic 16x16 tiles, "
                      f"median of {len(times)} run(s)",
            "config0": {"workload": "10k Gaussians uniform in [-1,1]^3, 256x256 view from (0,0,3), fovy 60",
                        "mrays_s": round(256 * 256 / med0 / 1e6, 4), "ms": round(med0 * 1e3, 3),
                        "host_sah_build_s": round(build0, 3), "threads": threads}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-multi", action="store_true", help="skip the multi-scene batch (256 views + 4M rays)")
    ap.add_argument("--no-hybrid", action="store_true", help="skip the hybrid (secondary-sphere) view")
    ap.add_argument("--no-viewer", action="store_true", help="skip the viewer frame loop")
    ap.add_argument("--no-order", action="store_true",
                    help="skip the raster-order and moving-camera legs of the plain view")
    ap.add_argument("--serve-scenes", type=int, default=16)
    ap.add_argument("--serve-window-ms", type=float, default=0.0,
                    help="how long a busy device waits to fill a batch (0: launch as soon as a slot is free)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--phases", action="store_true", help="add the per-phase cycle profile of the plain kernel")
    ap.add_argument("--rays", type=int, default=0,
                    help="also time kdgrt_render_rays over this many incoherent rays on the 1M scene (0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="batched leg with the default tuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    g, tree, build_s = scene_1m(args.gaussians)
    st = tree.stats()
    log(f"scene: {len(g)} Gaussians, {st['n_triangles']} proxy triangles, {st['n_nodes']} nodes, "
        f"{st['n_offsets']} leaf refs, depth {st['max_depth']}, build {build_s:.1f} s")
    aabb = tree.aabb
    dev = K.Device(0)
    info = dev.info()
    sc = dev.scene(tree, g)
    stream = dev.stream()
    W, H = 1920, 1080
    ci = view_1080p(aabb)
    out = dev.alloc(W * H * 16)

    for _ in range(args.warmup):
        dev.render(sc, ci, W, H, out, stream=stream)
    dev.sync()
    # Timed region: K launches on `stream`, HIP events on that stream.
    t0 = time.perf_counter()
    dev.timer_begin(stream)
    for _ in range(args.steps):
        dev.render(sc, ci, W, H, out, stream=stream)
    ms_total = dev.timer_end(stream)
    dev.sync()
    wall = time.perf_counter() - t0
    ms_launch = ms_total / args.steps
    rays = W * H
    value = rays / (ms_launch * 1e-3) / 1e6
    log(f"plain 1080p: {ms_launch:.3f} ms/launch, {value:.1f} Mrays/s (wall {wall:.3f} s for {args.steps})")

    dbg = dev.render_debug(sc, ci, W, H)
    cnt = dbg["counters"]
    B = algorithmic_bytes(cnt)
    mean_c = cnt.reshape(-1, K.N_COUNTERS).mean(0)
    log("mean counters/ray: " + ", ".join(f"{n}={v:.2f}" for n, v in zip(K.COUNTER_NAMES, mean_c)))
    # Compulsory bytes of one render call: k_view's records read once each
    # (touch maps of the debug kernel, same rays and tree) plus the cone
    # tables: k_gaussian_cone reads each Gaussian's 16-B bound sphere and
    # 24-B shape matrix and writes its 16-B cone record; k_cone_prep reads
    # each slot's 4-B device number and the Gaussian's record once and writes
    # the slot's 16-B cone record.
    touch = dev.render_touch(sc, ci, W, H)
    comp = K.compulsory_bytes(touch, tree.offsets, rays)
    n_slots, n_gauss = int(st["n_offsets"]) + 4, len(g)
    cone_prep = (16 + 24 + 16) * n_gauss + (4 + 16) * n_slots + 16 * n_gauss
    comp_launch = comp["this_layout"] + cone_prep
    achieved = comp_launch / (ms_launch * 1e-3) / 1e9
    traffic = load_profile("traffic.json", "plain_1080p_1M")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None if traffic is None else int(traffic["hbm_bytes_per_launch"]),
            "traffic_source": None if traffic is None else traffic.get("source"),
            "compulsory_bytes_per_launch": comp_launch,
            "compulsory_parts": {**comp["this_layout_parts"], "cone_prep": cone_prep},
            "compulsory_reference_layout_bytes": comp["reference_layout"],
            "logical_bytes_per_launch": B, "logical_bytes_per_ray": round(B / rays, 1),
            "logical_GB_s": round(B / (ms_launch * 1e-3) / 1e9, 1),
            "kernel": "k_gaussian_cone + k_cone_prep + k_view<7, false, 0, false, true> (one render call; k_cone_prep also sorts the "
                      "tile order)"}
    log(f"roofline: {roof}")

    phases = None
    if args.phases:
        import ctypes as C
        prof = dev.alloc(W * H * 8 * 8)
        K.check(K.lib().kdgrt_render_profile(dev.h, sc.h, C.byref(ci), C.c_int(W), C.c_int(H), out.ptr, prof.ptr,
                                             None))
        dev.sync()
        pr = prof.download(np.uint64, (H * W, 8)).astype(np.float64)
        prof.free()
        cyc = pr[:, :5]
        names = ["descend", "pretest", "exact_insert", "composite", "pop_loop"]
        tot = cyc.sum()
        phases = {n: round(float(cyc[:, i].sum() / tot), 4) for i, n in enumerate(names)}
        phases["mean_cycles_per_ray"] = round(float(cyc.sum(1).mean()), 1)
        phases["pretests_per_ray"] = round(float(pr[:, 5].mean()), 2)
        phases["exact_tests_per_ray"] = round(float(pr[:, 6].mean()), 2)
        log(f"phase shares: {phases}")

    copy = copy_bandwidth(dev, stream)
    log(f"hbm copy: {copy}")
    roof["frac_of_copy_ceiling"] = round(achieved / copy["GB_s"], 5)
    # VALU roofline: lane operations per SIMD-cycle of k_view against those of
    # k_valu_probe (every lane active, issue-limited), both from the counter
    # runs summarised in profiles/valu.json; the live probe rate is beside it.
    probe = dev.valu_probe(stream=stream)
    vprof = load_profile("valu.json", "plain_1080p_1M")
    valu = {"bound": "valu", "probe_live": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in probe.items()},
            "frac": None}
    if vprof is not None:
        valu.update({"lane_op_rate": round(vprof["lane_op_rate"], 5),
                     "probe_lane_op_rate": round(vprof["probe_lane_op_rate"], 5),
                     "frac": round(vprof["lane_op_rate"] / vprof["probe_lane_op_rate"], 4),
                     "issue_rate": round(vprof["issue_rate"], 4), "probe_issue_rate": round(vprof["probe_issue_rate"], 4),
                     "kernel_valu_instr_per_launch": vprof["sq_insts_valu_per_dispatch"],
                     "source": vprof.get("source")})
    log(f"valu: {valu}")
    # Texture-path ceiling, from the same counter runs: a 16-B vector load
    # instruction holds the path for at least 16 cycles per CU whatever the
    # number of active lanes (probe kernel), so reads x 16 / kernel cycles is
    # the share of the kernel that the load instructions alone fill.
    texture = None
    if vprof is not None and "td_floor_frac" in vprof:
        texture = {"bound": "texture path", "busy": round(vprof["td_busy"], 4),
                   "cycles_per_read": round(vprof["td_cycles_per_read"], 2), "floor_cycles_per_read": 16,
                   "reads_per_launch": vprof["vmem_reads_per_dispatch"],
                   "floor_frac": round(vprof["td_floor_frac"], 4), "source": vprof.get("source")}
    log(f"texture path: {texture}")
    result = {"metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(ms_launch, 4), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32", "data": "synthetic",
              "config": {"workload": f"plain one-view launch: {len(g)} clustered Gaussians, one 1920x1080 pinhole "
                                     f"view from radius 25, 1 ray/pixel, short stack 12, the view launched "
                                     f"repeatedly (the reference's render-test loop), tiles ordered by the "
                                     f"previous launch's tile durations",
                         "gaussians": len(g), "width": W, "height": H, "stack_depth": 12,
                         "tile_order": "frame-to-frame schedule (raster-order rate under tile_order)"},
              "device": info["name"], "roofline": roof, "valu": valu, "texture": texture, "hbm_copy": copy,
              **({"phases": phases} if phases else {}),
              "build": {"host_sah_build_s": round(build_s, 2), "threads": os.cpu_count(), "nodes": st["n_nodes"],
                        "leaf_refs": st["n_offsets"], "proxy_triangles": st["n_triangles"]}}

    if not args.no_order:
        result["tile_order"] = tile_order_legs(dev, sc, ci, aabb, W, H, out, stream, args.steps)

    if not args.no_hybrid:
        result["hybrid"] = hybrid_bench(dev, sc, ci, aabb, W, H, stream, args.steps, args.warmup)

    if not args.no_viewer:
        result["viewer"] = viewer_bench(dev, sc, aabb, W, H)

    if not args.no_batched:
        views = views_64(aabb)
        bufs, vlist = [], []
        total_rays = 0
        for vci, w, h in views:
            b = dev.alloc(w * h * 16)
            bufs.append(b)
            vlist.append(K.View(sc.h, vci, w, h, b.ptr))
            total_rays += w * h
        # plain: one launch per view, each waited for (the reference blocks per frame)
        for (vci, w, h), b in zip(views, bufs):
            dev.render(sc, vci, w, h, b, stream=stream)
        dev.sync(stream)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            for (vci, w, h), b in zip(views, bufs):
                dev.render(sc, vci, w, h, b, stream=stream)
                dev.sync(stream)
        plain_s = (time.perf_counter() - t0) / reps
        # the same launches queued back to back, one wait at the end
        t0 = time.perf_counter()
        for _ in range(reps):
            for (vci, w, h), b in zip(views, bufs):
                dev.render(sc, vci, w, h, b, stream=stream)
            dev.sync(stream)
        plain_nosync_s = (time.perf_counter() - t0) / reps
        # batched: autotune, then hipGraph replays of one persistent launch
        if args.no_autotune:
            tun, tms = K.default_tuning(dev), float("nan")
        else:
            tun, tms = K.autotune(dev, vlist, max_trials=20, reps=2)
        batch = K.Batch(dev, vlist, (), tun)
        batch.run(stream, use_graph=True)
        dev.sync(stream)
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.run(stream, use_graph=True)
            dev.sync(stream)
        batch_s = (time.perf_counter() - t0) / reps
        result["batched"] = {
            "views": 64, "rays": total_rays,
            "plain_per_view_mrays_s": round(total_rays / plain_s / 1e6, 2),
            "plain_per_view_nosync_mrays_s": round(total_rays / plain_nosync_s / 1e6, 2),
            "batched_graph_mrays_s": round(total_rays / batch_s / 1e6, 2),
            "speedup": round(plain_s / batch_s, 3), "speedup_vs_nosync": round(plain_nosync_s / batch_s, 3),
            "tuning": tun.as_dict(), "autotune_best_ms": round(tms, 3)}
        log(f"64 views ({total_rays} rays): plain {plain_s * 1e3:.2f} ms, batched {batch_s * 1e3:.2f} ms")
        batch.free()
        for b in bufs:
            b.free()

    if args.rays:
        result["incoherent_rays"] = rays_leg(dev, sc, aabb, args.rays, stream, args.steps, args.warmup)

    if not (args.no_serve and args.no_multi):
        tun = K.default_tuning(dev)
        if "batched" in result:
            tun = K.Tuning(**result["batched"]["tuning"])
        scenes, aabbs, counts, setup_s = scene_set(dev, args.serve_scenes)
        log(f"scene set: {len(scenes)} scenes, {int(counts.sum())} Gaussians, setup {setup_s:.1f} s")
        if not args.no_multi:
            m = min(8, len(scenes))
            result["multi_scene_batch"] = multi_scene_batch(dev, scenes[:m], aabbs[:m], counts[:m], tun)
        if not args.no_serve:
            result["serving"] = serve_bench(dev, scenes, aabbs, counts, tun, window_ms=args.serve_window_ms)
            result["serving"]["setup_s"] = round(setup_s, 1)
        for sc in scenes:
            sc.free()

    if not args.no_cpu:
        gi = K.invert_scales(g)
        result["cpu_baseline"] = cpu_baseline(tree, gi, ci, W, H, args.cpu_budget)
        log(f"cpu baseline: {result['cpu_baseline']}")

    out.free()
    sc.free()
    dev.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
