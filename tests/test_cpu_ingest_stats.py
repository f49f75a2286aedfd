"""Host products: NeRF transforms.json cameras (LoadCamera.hpp:25-62) and the
per-ray statistics files of a debug render (cudaRenderer.cu:3173-3400).

Cameras are checked against the matrix columns written into a synthetic
transforms file; the statistics files against values computed by hand from a
small counter array (averages over rays that blended, maxima over all rays,
totals, histogram rows, CSV matrices, the two uint32 files that both hold
hits_found)."""
import importlib
import json

import numpy as np
import pytest

K = importlib.import_module("3dgrt_kdtree_amd")


def test_nerf_cameras_closed_form(tmp_path):
    rng = np.random.default_rng(3)
    frames = []
    for _ in range(3):
        m = rng.normal(size=(4, 4)).round(6)
        m[3] = [0, 0, 0, 1]
        frames.append({"file_path": "./r_0", "rotation": 0.1, "transform_matrix": m.tolist()})
    p = tmp_path / "transforms_test.json"
    p.write_text(json.dumps({"camera_angle_x": 0.69, "frames": frames}, indent=2))
    cams = K.cameras_from_nerf_json(str(p), 1.5)
    assert len(cams) == 3
    for cam, f in zip(cams, frames):
        m = np.array(f["transform_matrix"], np.float64)
        assert np.array_equal(np.array(cam.u), m[:3, 0].astype(np.float32))
        assert np.array_equal(np.array(cam.v), m[:3, 1].astype(np.float32))
        assert np.array_equal(np.array(cam.n), m[:3, 2].astype(np.float32))
        assert np.array_equal(np.array(cam.pos), m[:3, 3].astype(np.float32))
        assert cam.fovy_deg == 39.6 and cam.near_c == 0.005 and cam.aspect == 1.5


@pytest.mark.parametrize("text", ['{"frames": [{"transform_matrix": [[1, 2], [3, 4]]}]}', '{"frames": 3}',
                                  '{"frames": [', '[1, 2, 3]'])
def test_nerf_cameras_rejects_malformed(tmp_path, text):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(K.KdgrtError):
        K.cameras_from_nerf_json(str(p), 1.0)


def test_stats_files_by_hand(tmp_path):
    W, H = 3, 2
    c = np.zeros((H, W, 8), np.int32)
    # (node, leaf, tests, hits, blends, max_sort) per ray; ray (1,0) never blends
    c[0, 0, :6] = (10, 2, 20, 3, 2, 2)
    c[0, 1, :6] = (50, 7, 90, 9, 0, 4)
    c[0, 2, :6] = (12, 2, 18, 1, 1, 1)
    c[1, 0, :6] = (10, 3, 30, 2, 2, 1)
    c[1, 1, :6] = (0, 0, 0, 0, 0, 0)
    c[1, 2, :6] = (7, 1, 5, 4, 3, 4)
    K.write_stats(c, W, H, tmp_path)
    blend = c[..., 4] != 0
    lines = (tmp_path / "renderStats.txt").read_text().splitlines()
    vals = {ln.split(":")[0].strip(): float(ln.split(":")[1]) for ln in lines if ":" in ln}
    names = ["node_visits", "leaf_visits", "intersection_tests", "hits_found", "blend_ops", "max_sort_size"]
    for k, nm in enumerate(names):
        assert vals["avg_" + nm] == pytest.approx(c[..., k][blend].mean(), abs=1e-6)
        assert vals["max_" + nm] == c[..., k].max()
        assert vals["tot_" + nm] == c[..., k][blend].sum()
    assert "avg_node_visits\t\t: 9.750000" in lines
    hist = (tmp_path / "render_stats.csv").read_text().splitlines()
    assert hist == ["Value,Node_Visits,Leaf_Visits", "1,0,1", "2,0,2", "3,0,1", "7,1,0", "10,2,0", "12,1,0"]
    assert (tmp_path / "heatmap_node_visits.csv").read_text() == "10,50,12\n10,0,7\n"
    assert (tmp_path / "heatmap_max_sort_size.csv").read_text() == "2,4,1\n1,0,4\n"
    for name in ("isectCnt_kd.bin", "blendCnt_kd.bin"):
        assert np.array_equal(np.fromfile(tmp_path / name, np.uint32), c[..., 3].reshape(-1).astype(np.uint32))


def _png_decode(data):
    """Minimal PNG reader for the writer's output: chunks with CRCs, one
    zlib stream, filter type 0 scanlines."""
    import struct
    import zlib
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) == crc, typ
        chunks.setdefault(typ, b"")
        chunks[typ] += body
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    ch = {2: 3, 6: 4}[ctype]
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * ch)
    assert (rows[:, 0] == 0).all() and depth == 8 and b"IEND" in chunks
    return rows[:, 1:].reshape(h, w, ch)


@pytest.mark.parametrize("shape", [(3, 5, 3), (37, 700, 4), (1, 1, 3), (130, 260, 3)])
def test_png_round_trip(tmp_path, shape):
    rng = np.random.default_rng(sum(shape))
    px = rng.integers(0, 256, shape, dtype=np.uint8)
    p = tmp_path / "f.png"
    K.write_png(p, px)
    assert np.array_equal(_png_decode(p.read_bytes()), px)


def test_png_rejects_bad_channels(tmp_path):
    with pytest.raises(K.KdgrtError):
        K.write_png(tmp_path / "x.png", np.zeros((2, 2, 2), np.uint8))
