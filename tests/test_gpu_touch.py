"""Touch maps (kdgrt_render_touch / kdgrt_render_rays_touch) against the
oracle's, and the compulsory-byte model built on them.

The maps say which records a launch reads; the roofline's compulsory bytes
count each once.  The device's node, leaf-slot, evaluated and blended maps
must equal the oracle's exactly; the exact-test map (a device notion: slots
that survive the conservative pre-test) must lie between the accepted slots
and the tested slots; every quad record loaded must be the child pair of a
visited inner node.
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


def _check(gpu, ref, nodes, label):
    print(f"[{label}] comparing touch maps", flush=True)
    assert np.array_equal(gpu["nodes"], ref["node"]), f"{label}: node map"
    assert np.array_equal(gpu["slots_pretested"], ref["slot"]), f"{label}: tested slots"
    assert np.array_equal(gpu["gaussians_evaluated"], ref["eval"]), f"{label}: evaluated Gaussians"
    assert np.array_equal(gpu["gaussians_blended"], ref["blend"]), f"{label}: blended Gaussians"
    ex = gpu["slots_tested"]
    assert not (ex & ~gpu["slots_pretested"]).any(), f"{label}: exact test outside the tested slots"
    assert not (ref["hit_slot"] & ~ex).any(), f"{label}: an accepted slot skipped the exact test"
    x = nodes["x"].astype(np.int64)
    inner = gpu["nodes"] & ((x & 7) != 3)
    allowed = np.zeros(len(nodes), bool)
    allowed[x[inner] >> 3] = True
    assert not (gpu["quads"] & ~allowed).any(), f"{label}: quad of an unvisited node"
    assert gpu["quads"].sum() > 0 and ex.sum() < gpu["slots_pretested"].sum()


def test_touch_view(dev):
    hs = S.clustered(100000)
    sc = dev.scene(hs.tree, hs.g)
    w, h = 320, 240
    ci = S.cam_orbit(hs.aabb, w, h)
    gpu = dev.render_touch(sc, ci, w, h)
    ref = O.touch(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, len(hs.g), ci=ci, w=w, h=h)
    _check(gpu, ref, hs.nodes, "view")
    assert np.array_equal(gpu["rgbt"], dev.render_image(sc, ci, w, h))
    # unique records never exceed the per-ray (algorithmic) count
    dbg = dev.render_debug(sc, ci, w, h)
    c = dbg["counters"].reshape(-1, K.N_COUNTERS).astype(np.int64)
    comp = K.compulsory_bytes(gpu, hs.offsets, w * h)
    alg = int(8 * (c[:, 0] + c[:, 1]).sum() + 52 * c[:, 2].sum() + 256 * c[:, 6].sum() + 16 * len(c))
    assert 0 < comp["reference_layout"] <= alg
    assert gpu["nodes"].sum() <= (c[:, 0] + c[:, 1]).sum()
    sc.free()


def test_touch_rays(dev):
    hs = S.clustered(100000)
    sc = dev.scene(hs.tree, hs.g)
    o, d = S.random_rays(hs.aabb, 30000, 23)
    gpu = dev.render_rays_touch(sc, o, d)
    ref = O.touch(hs.nodes, hs.offsets, hs.g_inv, hs.aabb, len(hs.g), origins=o, dirs=d)
    _check(gpu, ref, hs.nodes, "rays")
    assert np.array_equal(gpu["rgbt"], dev.render_rays(sc, o, d)["rgbt"])
    sc.free()


def test_valu_probe_counts(dev):
    """The calibration kernel runs and reports its instruction count."""
    st = dev.stream()
    r = dev.valu_probe(1 << 12, stream=st)
    assert r["fma_per_wave"] == 8 << 12 and r["waves"] >= 8 * 4 * 256 // 8
    assert r["valu_instr_per_s"] > 1e11
