"""Host prerequisites: proxy meshing, SAH builder, .kdt format, validation, PLY.

The fast builder is pinned against ref_builder_literal.py, a statement-by-
statement restatement of the reference build (edges in qsort order, per-
position Gaussian-ID maps, triangle clipping in float32): whole trees must be
byte-identical.
"""
import importlib
import math
import os
import struct

import numpy as np
import pytest

import ref_builder_literal as RB

K = importlib.import_module("3dgrt_kdtree_amd")


@pytest.mark.parametrize("kind,n,seed,posscale", [(0, 60, 3, 0.2), (1, 200, 4, 1.0), (0, 300, 5, 1.0),
                                                  (0, 250, 13, 0.05)])
def test_builder_matches_literal_restatement(kind, n, seed, posscale):
    g = K.generate_gaussians(kind, n, seed)
    g["pos"] *= posscale
    tree = K.Tree.build(g)
    xyz, gid, aabb = K.proxy_mesh(g)
    nodes, offs = RB.build(xyz, gid, aabb)
    fast = tree.nodes
    assert len(fast) == len(nodes)
    assert [(int(a["x"]), int(a["y"])) for a in fast] == nodes
    assert list(tree.offsets) == offs
    assert np.array_equal(tree.aabb, aabb)


def test_builder_sah_mode0_matches_literal():
    g = K.generate_gaussians(0, 200, 21)
    p = K.default_build_params()
    p.sah_mode = 0
    tree = K.Tree.build(g, p)
    xyz, gid, aabb = K.proxy_mesh(g)
    P = dict(travl=np.float32(1), isect=np.float32(5), bonus=np.float32(0.9), max_level=256, min_prims=16,
             force=256, sah_mode=0)
    nodes, offs = RB.build(xyz, gid, aabb, P)
    assert [(int(a["x"]), int(a["y"])) for a in tree.nodes] == nodes and list(tree.offsets) == offs


def test_builder_thread_count_invariant():
    g = K.generate_gaussians(1, 20000, 31)
    a = K.Tree.build(g, threads=1)
    b = K.Tree.build(g, threads=8)
    assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.offsets, b.offsets)


def quat_rotate(q, v):
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return v @ R.T


def test_proxy_mesh_closed_form_and_cull():
    g = K.generate_gaussians(0, 500, 7)
    g["opacity"][:5] = [0.01, 0.0112, 0.003, 0.0113, 0.5]  # first three culled
    xyz, gid, aabb = K.proxy_mesh(g)
    kept = ~((g["opacity"] < np.float32(0.0113)) | (g["opacity"] < np.float32(1 / 255.0)))
    assert len(gid) == 20 * kept.sum()
    assert set(np.unique(gid)) == set(np.nonzero(kept)[0])
    phi = 1.618033988749895
    ico = np.array([[-1, phi, 0], [1, phi, 0], [0, 1, -phi], [-phi, 0, -1], [-phi, 0, 1], [0, 1, phi], [phi, 0, 1],
                    [0, -1, phi], [-1, -phi, 0], [0, -1, -phi], [phi, 0, -1], [1, -phi, 0]])
    for i in np.nonzero(kept)[0][:50]:
        o = float(g["opacity"][i])
        k = math.sqrt(2 * math.log(o / 0.0113)) * 0.5257311121191336 if o / 0.0113 > 1 else 0.0
        verts = quat_rotate(g["rot"][i].astype(np.float64), ico * (g["scale"][i].astype(np.float64) * k)) + g["pos"][i]
        tri = xyz[gid == i].reshape(-1, 3)
        # every proxy vertex lies on the k_iso ellipsoid and is one of the 12
        d = np.abs(tri[:, None, :] - verts[None, :, :]).max(-1).min(-1)
        assert d.max() < 1e-5 * (1 + np.abs(verts).max())
    allv = xyz.reshape(-1, 3)
    assert np.array_equal(aabb[:3], allv.min(0)) and np.array_equal(aabb[3:], allv.max(0))


def _leaf_of_point(nodes, p):
    i = 0
    while (int(nodes[i]["x"]) & 7) != 3:
        x = int(nodes[i]["x"])
        axis, c = x & 3, x >> 3
        split = np.array([nodes[i]["y"]], np.uint32).view(np.float32)[0]
        i = c if p[axis] < split else c + 1
    return i


def test_builder_invariants_and_coverage():
    g = K.generate_gaussians(1, 30000, 41)
    tree = K.Tree.build(g)
    tree.validate(len(g))
    nodes, offs = tree.nodes, tree.offsets
    st = tree.stats()
    assert st["n_nodes"] == 2 * st["n_leaves"] - 1
    for i in range(len(nodes)):
        x = int(nodes[i]["x"])
        if (x & 7) == 3:
            cnt, off = x >> 3, int(nodes[i]["y"])
            ids = offs[off:off + cnt]
            assert np.all(np.diff(ids.astype(np.int64)) > 0)  # sorted unique (std::set order)
        else:
            assert (x >> 3) > i and (x & 3) < 3
    # Coverage: the centroid of every proxy triangle lies in a leaf that lists its Gaussian.
    xyz, gid, _ = K.proxy_mesh(g)
    rng = np.random.default_rng(0)
    for t in rng.choice(len(gid), 3000, replace=False):
        c = xyz[t].astype(np.float64).mean(0).astype(np.float32)
        li = _leaf_of_point(nodes, c)
        x = int(nodes[li]["x"])
        ids = offs[int(nodes[li]["y"]):int(nodes[li]["y"]) + (x >> 3)]
        assert gid[t] in ids


def test_kdt_roundtrip_and_layout(tmp_path):
    g = K.generate_gaussians(0, 3000, 5)
    tree = K.Tree.build(g)
    p = str(tmp_path / "t.kdt")
    tree.write_kdt(p)
    raw = open(p, "rb").read()
    ver, = struct.unpack("<f", raw[:4])
    assert abs(ver - 0.1) < 1e-7
    aabb = np.frombuffer(raw[4:28], np.float32)
    assert np.array_equal(aabb, tree.aabb)
    nn, = struct.unpack("<i", raw[28:32])
    assert nn == len(tree.nodes)
    assert len(raw) == 32 + 8 * nn + 4 + 4 * len(tree.offsets)
    t2 = K.Tree.read_kdt(p)
    assert np.array_equal(t2.nodes, tree.nodes) and np.array_equal(t2.offsets, tree.offsets)
    assert np.array_equal(t2.aabb, tree.aabb)


def test_kdt_corrupt_files_rejected(tmp_path):
    g = K.generate_gaussians(0, 500, 5)
    tree = K.Tree.build(g)
    p = str(tmp_path / "t.kdt")
    tree.write_kdt(p)
    raw = bytearray(open(p, "rb").read())
    cases = {"truncated": raw[:-3], "empty": b"", "header": raw[:20]}
    bad_child = bytearray(raw)
    struct.pack_into("<I", bad_child, 32, (10 ** 6 << 3) | 0)  # root child far out of range
    cases["child"] = bad_child
    bad_count = bytearray(raw)
    struct.pack_into("<i", bad_count, 28, -5)
    cases["count"] = bad_count
    for name, data in cases.items():
        q = str(tmp_path / f"{name}.kdt")
        open(q, "wb").write(bytes(data))
        with pytest.raises(K.KdgrtError):
            K.Tree.read_kdt(q)
    with pytest.raises(K.KdgrtError):
        K.Tree.read_kdt(str(tmp_path / "missing.kdt"))


def test_validate_rejects_bad_trees():
    nodes = np.zeros(3, K.NODE_DTYPE)
    nodes[0] = ((1 << 3) | 0, np.float32(0).view(np.uint32))
    nodes[1] = ((1 << 3) | 3, 0)
    nodes[2] = ((1 << 3) | 3, 1)
    aabb = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    t = K.Tree.from_arrays(nodes, np.array([0, 1], np.uint32), aabb)
    t.validate(2)
    with pytest.raises(K.KdgrtError):
        t.validate(1)  # GID 1 >= 1
    loop = nodes.copy()
    loop[0] = ((0 << 3) | 0, 0)  # child index not after parent
    with pytest.raises(K.KdgrtError):
        K.Tree.from_arrays(loop, np.array([0, 1], np.uint32), aabb)
    rng_bad = nodes.copy()
    rng_bad[2] = ((5 << 3) | 3, 1)  # offset range past the list
    with pytest.raises(K.KdgrtError):
        K.Tree.from_arrays(rng_bad, np.array([0, 1], np.uint32), aabb)


def write_ply(path, g_raw):
    names = ["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)] + \
            [f"f_rest_{i}" for i in range(45)] + ["opacity"] + [f"scale_{i}" for i in range(3)] + \
            [f"rot_{i}" for i in range(4)]
    hdr = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % len(g_raw)
    hdr += "".join(f"property float {n}\n" for n in names) + "end_header\n"
    with open(path, "wb") as f:
        f.write(hdr.encode())
        f.write(g_raw.astype("<f4").tobytes())
    return names


def test_ply_reader_activation(tmp_path):
    rng = np.random.default_rng(3)
    n = 50
    raw = rng.normal(size=(n, 62)).astype(np.float32)
    raw[:, 54] = 0.0  # rot_0 .. zero quaternion in row 0 below
    raw[0, 58:62] = 0.0
    names = write_ply(str(tmp_path / "a.ply"), raw)
    col = {nm: i for i, nm in enumerate(names)}
    h = K.C.c_uint64()
    K.check(K.lib().kdgrt_ply_count(str(tmp_path / "a.ply").encode(), K.C.byref(h)))
    assert h.value == n
    g = np.zeros(n, K.GAUSSIAN_DTYPE)
    K.check(K.lib().kdgrt_ply_read(str(tmp_path / "a.ply").encode(), K._p(g), K.C.c_uint64(n), K.C.byref(h)))
    assert np.array_equal(g["pos"], raw[:, [col["x"], col["y"], col["z"]]])
    assert np.allclose(g["opacity"], 1 / (1 + np.exp(-raw[:, col["opacity"]].astype(np.float64))), rtol=1e-6)
    sc = raw[:, [col["scale_0"], col["scale_1"], col["scale_2"]]].astype(np.float64)
    assert np.allclose(g["scale"], np.exp(sc), rtol=1e-6)
    q = raw[:, [col[f"rot_{i}"] for i in range(4)]].astype(np.float64)
    norm = np.linalg.norm(q, axis=1)
    ok = norm ** 2 > 1e-5
    assert np.allclose(g["rot"][ok], q[ok] / norm[ok, None], atol=1e-6)
    assert np.array_equal(g["rot"][0], np.zeros(4, np.float32))  # not normalised (len^2 <= 1e-5)
    assert np.array_equal(g["f_rest"], raw[:, col["f_rest_0"]:col["f_rest_44"] + 1])


def test_generators_deterministic_and_distributed():
    a = K.generate_gaussians(1, 20000, 1001)
    b = K.generate_gaussians(1, 20000, 1001)
    assert a.tobytes() == b.tobytes()
    c = K.generate_gaussians(0, 20000, 1000)
    assert np.all(np.abs(c["pos"]) <= 1)
    ls = np.log(c["scale"].astype(np.float64))
    assert abs(ls.mean() - math.log(0.01)) < 0.02 and abs(ls.std() - 0.5) < 0.02
    assert np.allclose(np.linalg.norm(c["rot"], axis=1), 1, atol=1e-6)
    logit = np.log(c["opacity"] / (1 - c["opacity"]))
    assert abs(logit.std() - 1.0) < 0.03
    rgb = c["f_dc"] * 0.28209479177387814 + 0.5
    assert rgb.min() >= -1e-6 and rgb.max() <= 1 + 1e-6
    la = np.log(a["scale"].astype(np.float64))
    assert abs(la.mean() - math.log(0.02)) < 0.02 and abs(la.std() - 0.7) < 0.02
    bg = np.all(np.abs(a["pos"]) <= 10, axis=1).mean()
    assert bg > 0.5
