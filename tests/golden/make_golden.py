"""Writes the regression fixtures of tests/golden/ (run from the repository
root: python tests/golden/make_golden.py).

The reference ships no test vectors and its builder and renderer cannot be
built here, so these fixtures are produced by this repository's SAH builder
(itself pinned byte-for-byte against the literal restatement in
tests/ref_builder_literal.py) and FP32 oracle (pinned by the closed-form
tests of tests/test_cpu_oracle.py).  They anchor both against regressions:
  golden_3k.kdt        tree of 3000 uniform Gaussians (generator kind 0, seed 4242)
  golden_3k_view.npz   64x64 view from (0, 0, 3): RGBT, 8 counters, FNV hash,
                        first 16 composited IDs per ray
  golden_3k_rays.npz   512 incoherent rays (seed 4243): origins, directions,
                        RGBT, counters, hash
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
import scenes as S  # noqa: E402

K = importlib.import_module("3dgrt_kdtree_amd")

N, SEED, W, H, NRAYS = 3000, 4242, 64, 64, 512


def golden_scene():
    g = K.generate_gaussians(0, N, SEED)
    return g, K.Tree.build(g)


def golden_camera():
    return S.cam_c0(W, H)


def golden_rays(aabb):
    return S.random_rays(aabb, NRAYS, SEED + 1)


def main():
    g, tree = golden_scene()
    tree.write_kdt(os.path.join(HERE, "golden_3k.kdt"))
    gi = K.invert_scales(g)
    r = O.render(tree.nodes, tree.offsets, gi, tree.aabb, golden_camera(), W, H, seq_cap=16)
    np.savez_compressed(os.path.join(HERE, "golden_3k_view.npz"), rgbt=r["rgbt"], counters=r["counters"],
                        hash=r["hash"], seq=r["seq"])
    o, d = golden_rays(tree.aabb)
    rr = O.render_rays(tree.nodes, tree.offsets, gi, tree.aabb, o, d)
    np.savez_compressed(os.path.join(HERE, "golden_3k_rays.npz"), origins=o, dirs=d, rgbt=rr["rgbt"],
                        counters=rr["counters"], hash=rr["hash"])
    print("wrote fixtures:", sorted(f for f in os.listdir(HERE) if f.startswith("golden_3k")))


if __name__ == "__main__":
    main()
