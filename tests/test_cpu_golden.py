"""Regression fixtures (tests/golden/, written by tests/golden/make_golden.py):
the SAH builder must reproduce golden_3k.kdt byte for byte and the FP32
oracle the stored images, counters and composited-ID hashes bit for bit."""
import importlib
import os
import sys

import numpy as np

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402

K = importlib.import_module("3dgrt_kdtree_amd")
GOLD = os.path.join(HERE, "golden")


def test_golden_tree_bytes(tmp_path):
    g, tree = MG.golden_scene()
    out = tmp_path / "t.kdt"
    tree.write_kdt(str(out))
    assert out.read_bytes() == open(os.path.join(GOLD, "golden_3k.kdt"), "rb").read()
    st = tree.stats()
    assert st["n_nodes"] > 10 and st["n_offsets"] > 3000  # several leaves, clipped references


def test_golden_oracle_view_and_rays():
    g, tree = MG.golden_scene()
    gi = K.invert_scales(g)
    ref = np.load(os.path.join(GOLD, "golden_3k_view.npz"))
    r = O.render(tree.nodes, tree.offsets, gi, tree.aabb, MG.golden_camera(), MG.W, MG.H, seq_cap=16)
    for k in ("rgbt", "counters", "hash", "seq"):
        assert np.array_equal(r[k], ref[k]), k
    assert ref["counters"][..., 4].sum() > 0 and ref["counters"][..., 1].max() > 1  # blends, several leaves
    rr_ref = np.load(os.path.join(GOLD, "golden_3k_rays.npz"))
    o, d = MG.golden_rays(tree.aabb)
    assert np.array_equal(o, rr_ref["origins"]) and np.array_equal(d, rr_ref["dirs"])
    rr = O.render_rays(tree.nodes, tree.offsets, gi, tree.aabb, o, d)
    for k in ("rgbt", "counters", "hash"):
        assert np.array_equal(rr[k], rr_ref[k]), k
