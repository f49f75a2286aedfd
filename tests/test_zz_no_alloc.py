"""kdgrt_render_rays_noalloc (ray lists with the sort scratch in a static
device array) against kdgrt_render_rays, through the C ABI.

Bar: bit-identical RGBT per ray.  Calls in a row with shrinking inputs (the
static array keeps the previous call's contents), a list too short to be
sorted, the largest list the array holds, and the capacity error above it.
"""
import ctypes as C
import importlib

import numpy as np
import pytest

import oracle_lib as O
import scenes as S

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 21  # the static array's capacity (include/kdgrt.h)


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _render(dev, sc, o, d, noalloc, stream):
    n = len(o)
    ob, db, out = dev.upload(o), dev.upload(d), dev.alloc(n * 16)
    out.upload(np.full((n, 4), np.nan, np.float32))
    fn = K.lib().kdgrt_render_rays_noalloc if noalloc else K.lib().kdgrt_render_rays
    K.check(fn(dev.h, sc.h, ob.ptr, db.ptr, C.c_uint64(n), C.c_int(12), out.ptr, stream))
    dev.sync(stream)
    r = out.download(np.float32, (n, 4))
    for b in (ob, db, out):
        b.free()
    return r


def test_noalloc_equals_render_rays(dev):
    hs = S.clustered(100000)
    sc = dev.scene(hs.tree, hs.g)
    stream = dev.stream()
    for n, seed in ((300000, 41), (90000, 42), (1000, 43), (MAX_RAYS, 44), (70000, 45)):
        print(f"[noalloc] {n} rays, seed {seed}", flush=True)
        o, d = S.random_rays(hs.aabb, n, seed)
        a = _render(dev, sc, o, d, True, stream)
        b = _render(dev, sc, o, d, False, stream)
        assert not np.isnan(a).any()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{n} rays: differs from kdgrt_render_rays"
        if n <= 100000:
            ref = O.render_rays(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, o[:5000], d[:5000])
            assert np.array_equal(a[:5000], ref["rgbt"])
    sc.free()


def test_noalloc_capacity_error(dev):
    hs = S.uniform_10k()
    sc = dev.scene(hs.tree, hs.g)
    dummy = dev.alloc(256)
    st = K.lib().kdgrt_render_rays_noalloc(dev.h, sc.h, dummy.ptr, dummy.ptr, C.c_uint64(MAX_RAYS + 1), C.c_int(12),
                                           dummy.ptr, None)
    assert st == 6  # KDGRT_ERR_CAPACITY
    assert b"static scratch" in K.lib().kdgrt_last_error()
    assert K.lib().kdgrt_render_rays_noalloc(dev.h, sc.h, None, None, C.c_uint64(5), C.c_int(12), None, None) != 0
    assert K.lib().kdgrt_render_rays_noalloc(dev.h, sc.h, dummy.ptr, dummy.ptr, C.c_uint64(1), C.c_int(0),
                                             dummy.ptr, None) != 0
    dummy.free()
    sc.free()
