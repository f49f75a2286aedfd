"""Frame loop of the interactive viewer (kdgrt_viewer_*) on the GPU.

The reference renders each frame into an OpenGL display buffer mapped for
CUDA (Kd-treeConverterMain.cpp:226-251, 3514-3608); here frames go through a
ring of slots to pinned host memory.  Bar:
  * every delivered frame equals, bit for bit, the plain launch of the same
    camera (kdgrt_render, float3 framebuffer) -- and for the 8-bit formats
    the device conversion of that framebuffer (kdgrt_framebuffer_to_u8,
    rows top to bottom) -- along a camera path of orbit steps, key moves and
    drags; one frame also equals the FP32 oracle;
  * frames come back in submission order with the camera they were
    rendered with;
  * a slot is reused only after release: with every slot held, submit
    returns KDGRT_ERR_NOT_READY and no held frame changes;
  * the FPS window averages 1000 / render_ms over the frames in it.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

W, H = 320, 240


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def scene10k(dev):
    hs = S.uniform_10k()
    sc = dev.scene(hs.tree, hs.g)
    yield hs, sc
    sc.free()


def camera_path(n):
    """Start 3 units out, then orbit about the scene's y axis with key moves
    and drags on some frames."""
    cam = K.look_at((1.2, 0.9, 2.6), (0.0, 0.0, 0.0), 60.0, W / H, 0.005)
    out = []
    for i in range(n):
        K.camera_orbit(cam, 0.8 / 60.0, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
        if i % 3 == 1:
            K.camera_move(cam, K.KEY_W | K.KEY_D, 0.5, 1.0 / 60.0)
        if i % 4 == 2:
            K.camera_drag(cam, 7, -3, 0.1)
        out.append(K.camera_copy(cam))
    return out


def expected(dev, sc, cam, fmt):
    ci = K.camera_make(cam, W, H)
    fb = dev.alloc(W * H * 12)
    try:
        dev.render(sc, ci, W, H, fb, fmt=K.OUT_RGB)
        dev.sync()
        if fmt == K.FRAME_RGB_F32:
            return fb.download(np.float32, (H, W, 3))
        return dev.framebuffer_to_u8(fb, W, H, 3, 4 if fmt == K.FRAME_RGBA8 else 3, flip_rows=True)
    finally:
        fb.free()


@pytest.mark.parametrize("fmt", [K.FRAME_RGB_F32, K.FRAME_RGBA8, K.FRAME_RGB8])
def test_frames_equal_plain_render(dev, scene10k, fmt):
    hs, sc = scene10k
    cams = camera_path(10)
    v = K.Viewer(dev, sc, W, H, fmt=fmt, slots=3, window=(2, 8))
    print(f"[viewer] format {fmt}: 10 frames", flush=True)
    got = []
    pending = 0
    for cam in cams:
        v.submit(cam)
        pending += 1
        if pending == 2:                  # keep one frame in flight behind the one taken
            f, px = v.acquire()
            got.append((f.frame_id, f.render_ms, K.camera_copy(f.camera), px))
            v.release(f.frame_id)
            pending -= 1
    while pending:
        f, px = v.acquire()
        got.append((f.frame_id, f.render_ms, K.camera_copy(f.camera), px))
        v.release(f.frame_id)
        pending -= 1
    assert [g[0] for g in got] == list(range(10))
    for (fid, ms, fcam, px), cam in zip(got, cams):
        assert bytes(fcam) == bytes(cam)
        ref = expected(dev, sc, cam, fmt)
        assert px.shape == ref.shape and px.dtype == ref.dtype
        assert np.array_equal(px.view(np.uint8), ref.view(np.uint8)), f"frame {fid} differs from the plain render"
        assert ms > 0
    st = v.stats()
    assert st["submitted"] == 10 and st["delivered"] == 10 and st["window_frames"] == 6
    want = np.mean([1000.0 / np.float32(g[1]) for g in got[2:8]])
    assert abs(st["window_avg_fps"] - want) <= 1e-3 * want
    v.close()


def test_frame_equals_oracle(dev, scene10k):
    hs, sc = scene10k
    cam = camera_path(5)[-1]
    v = K.Viewer(dev, sc, W, H, fmt=K.FRAME_RGB_F32, slots=2)
    v.submit(cam)
    f, px = v.acquire()
    v.release(f.frame_id)
    v.close()
    ref = O.render(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, K.camera_make(cam, W, H), W, H)["rgbt"]
    assert np.array_equal(px.view(np.uint32), np.ascontiguousarray(ref[..., :3]).view(np.uint32))


def test_held_slots_are_not_reused(dev, scene10k):
    hs, sc = scene10k
    cams = camera_path(6)
    v = K.Viewer(dev, sc, W, H, fmt=K.FRAME_RGBA8, slots=3)
    print("[viewer] ring of 3 slots", flush=True)
    for cam in cams[:3]:
        v.submit(cam)
    held = []
    for _ in range(3):
        f, px = v.acquire(copy=False)
        held.append((f.frame_id, px, px.copy()))
    with pytest.raises(K.KdgrtError) as e:
        v.submit(cams[3], timeout_ms=0)
    assert e.value.status == 7                                    # KDGRT_ERR_NOT_READY
    with pytest.raises(K.KdgrtError):
        v.acquire(timeout_ms=0)                                   # nothing pending
    v.release(held[0][0])
    v.submit(cams[3], timeout_ms=0)                               # the released slot takes it
    f, px = v.acquire()
    assert f.frame_id == 3
    assert np.array_equal(px, expected(dev, sc, cams[3], K.FRAME_RGBA8))
    for fid, view, snap in held[1:]:
        assert np.array_equal(view, snap), f"held frame {fid} was overwritten"
    with pytest.raises(K.KdgrtError):
        v.release(held[0][0])                                     # already released
    v.close()
