"""Frame-to-frame tile schedule of the plain view launch (kdgrt_render).

The launch order of the tiles changes no pixel: frames rendered with the
schedule on (later launches of a nearly unchanged view take the tiles that
ran longest first) equal frames rendered in raster order bit for bit, over
repeated views and small camera moves (the order is followed), a camera jump
and size changes (it is not), views small enough for 4x4-pixel tiles, and
the RGB output format.  Output buffers are filled with NaN before every launch
so that a tile left out by an order would show.
"""
import importlib

import numpy as np
import pytest

import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _frame(dev, sc, ci, w, h, fmt=K.OUT_RGBT):
    ch = 4 if fmt == K.OUT_RGBT else 3
    buf = dev.upload(np.full((h, w, ch), np.nan, np.float32))
    dev.render(sc, ci, w, h, buf, fmt=fmt)
    dev.sync()
    img = buf.download(np.float32, (h, w, ch))
    buf.free()
    return img


def _orbit(aabb, w, h, yaw_deg):
    a = np.radians(yaw_deg)
    d = np.array([0.48, 0.36, 0.8])
    d = np.array([d[0] * np.cos(a) + d[2] * np.sin(a), d[1], -d[0] * np.sin(a) + d[2] * np.cos(a)])
    return S.cam_orbit(aabb, w, h, direction=d)


def test_schedule_changes_no_pixel(dev):
    hs = S.clustered(100_000, seed=1011)
    on = dev.scene(hs.tree, hs.g)
    off = dev.scene(hs.tree, hs.g)
    off.set_tile_schedule(False)
    aabb = hs.aabb
    seq = [(0.0, 640, 480)] * 3 + [(0.2, 640, 480), (0.35, 640, 480)]  # repeats, then moves within 0.5 degrees
    seq += [(40.0, 640, 480), (40.0, 640, 480)]                        # a jump, then a repeat
    seq += [(40.0, 1280, 720)] * 2 + [(40.0, 160, 120)] * 3            # size changes; 160x120 runs as 4x4 tiles
    seq += [(0.0, 640, 480)] * 2 + [(0.0, 1001, 333)] * 2               # back; a size with partial tiles
    for k, (yaw, w, h) in enumerate(seq):
        ci = _orbit(aabb, w, h, yaw)
        print(f"[schedule] frame {k}: yaw {yaw} deg, {w}x{h}", flush=True)
        a = _frame(dev, on, ci, w, h)
        b = _frame(dev, off, ci, w, h)
        assert not np.isnan(a).any(), f"frame {k}: pixels left unwritten"
        assert np.array_equal(a, b), f"frame {k}: scheduled and raster frames differ"
    ci = _orbit(aabb, 640, 480, 0.0)
    for _ in range(2):
        assert np.array_equal(_frame(dev, on, ci, 640, 480, K.OUT_RGB), _frame(dev, off, ci, 640, 480, K.OUT_RGB))
    # the debug kernel (raster order, counters) writes the same image
    assert np.array_equal(_frame(dev, on, ci, 640, 480), dev.render_debug(on, ci, 640, 480)["rgbt"])
    on.free()
    off.free()


def test_schedule_toggle_and_repeat_full_size(dev):
    """The 1M-Gaussian 1080p view, five launches with the schedule: all equal
    to the raster launch; turning the schedule off and on again keeps it so."""
    hs = S.clustered(1_000_000)
    sc = dev.scene(hs.tree, hs.g)
    ci = S.cam_orbit(hs.aabb, 1920, 1080)
    sc.set_tile_schedule(False)
    ref = _frame(dev, sc, ci, 1920, 1080)
    sc.set_tile_schedule(True)
    for k in range(5):
        print(f"[schedule 1M] launch {k}", flush=True)
        assert np.array_equal(_frame(dev, sc, ci, 1920, 1080), ref), f"launch {k} differs"
    sc.set_tile_schedule(False)
    assert np.array_equal(_frame(dev, sc, ci, 1920, 1080), ref)
    sc.set_tile_schedule(True)
    assert np.array_equal(_frame(dev, sc, ci, 1920, 1080), ref)
    sc.free()


def test_schedule_hybrid_and_interleaved_kernels(dev):
    """The hybrid kernel follows the same schedule: hybrid frames with the
    schedule equal raster-order hybrid frames, also when plain and hybrid
    launches of one view alternate (each kernel keeps its own durations)."""
    hs = S.clustered(100_000, seed=1012)
    on = dev.scene(hs.tree, hs.g)
    off = dev.scene(hs.tree, hs.g)
    off.set_tile_schedule(False)
    sph = S.spheres_in_view(hs.aabb)
    w, h = 640, 480

    def hyb(sc, ci):
        buf = dev.upload(np.full((h, w, 3), np.nan, np.float32))
        dev.render_hybrid(sc, ci, w, h, spheres=sph, out=buf)
        dev.sync()
        img = buf.download(np.float32, (h, w, 3))
        buf.free()
        return img

    for k, (yaw, kind) in enumerate([(0.0, "h"), (0.0, "h"), (0.0, "h"), (0.2, "h"), (0.2, "p"), (0.2, "h"),
                                     (0.2, "p"), (0.2, "p"), (30.0, "h"), (30.0, "h")]):
        ci = _orbit(hs.aabb, w, h, yaw)
        print(f"[schedule hybrid] frame {k}: yaw {yaw}, kernel {kind}", flush=True)
        if kind == "h":
            a, b = hyb(on, ci), hyb(off, ci)
        else:
            a, b = _frame(dev, on, ci, w, h), _frame(dev, off, ci, w, h)
        assert not np.isnan(a).any(), f"frame {k}: pixels left unwritten"
        assert np.array_equal(a, b), f"frame {k}: scheduled and raster frames differ"
    on.free()
    off.free()


def test_schedule_across_streams(dev):
    """Launches of one scene alternate between two streams without host
    waits in between: each launch's tile order comes from the durations the
    other stream's launch recorded, ordered by the scene's event.  Every
    frame equals the raster-order frame."""
    hs = S.clustered(100_000, seed=1013)
    sc = dev.scene(hs.tree, hs.g)
    w, h = 640, 480
    ci = _orbit(hs.aabb, w, h, 0.0)
    sc.set_tile_schedule(False)
    ref = _frame(dev, sc, ci, w, h)
    sc.set_tile_schedule(True)
    streams = [dev.stream(), dev.stream()]
    bufs = [dev.upload(np.full((h, w, 4), np.nan, np.float32)) for _ in range(8)]
    for k, b in enumerate(bufs):
        dev.render(sc, ci, w, h, b, stream=streams[k % 2])
    dev.sync()
    for k, b in enumerate(bufs):
        img = b.download(np.float32, (h, w, 4))
        assert np.array_equal(img, ref), f"launch {k} (stream {k % 2}) differs"
        b.free()
    for st in streams:
        K.check(K.lib().kdgrt_stream_destroy(dev.h, st))
    sc.free()


@pytest.mark.parametrize("wh", [(1, 1), (13, 7), (65, 33), (8, 200), (300, 9), (2000, 24)])
def test_schedule_partial_tiles(dev, wh):
    """Sizes that leave partial tiles, down to one pixel: the first launch
    (ordered by centre density) and a repeat (ordered by durations) equal the
    raster-order debug kernel's image."""
    w, h = wh
    hs = S.uniform_10k()
    ci = S.cam_c0(w, h)
    sc = dev.scene(hs.tree, hs.g)
    ref = dev.render_debug(sc, ci, w, h)["rgbt"]
    for k in range(2):
        print(f"[schedule] {w}x{h} launch {k}", flush=True)
        img = _frame(dev, sc, ci, w, h)
        assert not np.isnan(img).any(), f"{w}x{h} launch {k}: pixels left unwritten"
        assert np.array_equal(img, ref), f"{w}x{h} launch {k} differs"
    sc.free()
