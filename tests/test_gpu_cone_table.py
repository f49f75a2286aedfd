"""The cone table a plain view pre-tests (kdgrt_view_cone_table) against a
CPU restatement.

Output parity cannot see every fault of this table: a record that accepts
everything only slows the view down.  So the table itself is checked:

* every leaf slot's record holds a = mu - eye of the slot's Gaussian
  (offsets[i], the caller's numbering) in f32, bit for bit: the gather of
  k_cone_prep through the device numbering, its per-wave chunks and its
  tail, at scene sizes from a few thousand slots to the 9.4M of the 1M scene;
* the four padding records are zero;
* q is never above the larger of the two bounds of k_gaussian_cone restated
  in float64 (projected ellipsoid and bound sphere, as in
  tests/test_cpu_cone_bound.py, which pins both against the oracle's exact
  test) with the stored kappa (rounded up to f16), and never below the
  ellipsoid bound, up to rounding: the device's bound-sphere radius (also
  rounded up to f16) can only loosen its bound, and q is rounded down to f32
  once;
* a Gaussian without a bound (the eye inside its slab) has q = -FLT_MAX.
"""
import importlib

import numpy as np
import pytest

import scenes as S
from test_cpu_cone_bound import q_ellipse, q_sphere, shape

K = importlib.import_module("3dgrt_kdtree_amd")

pytestmark = pytest.mark.gpu

NO_BOUND = np.float32(-3.402823466e+38)


def f16_up(x):
    """x rounded up to a half-precision value: the bound-sphere record
    stores its radius and kappa so."""
    h = np.float16(x)
    return float(np.nextafter(h, np.float16(np.inf)) if float(h) < x else h)


@pytest.fixture(scope="module")
def dev():
    d = K.Device(0)
    yield d
    d.close()


def _check(dev, hs, eye, n_q, seed):
    sc = dev.scene(hs.tree, hs.g)
    e = np.asarray(eye, np.float32)
    print(f"[cone table] {len(hs.g)} Gaussians, {len(hs.offsets)} slots, eye {tuple(eye)}", flush=True)
    tab = dev.view_cone_table(sc, e)
    sc.free()
    ids = np.asarray(hs.offsets, np.int64)
    n = len(ids)
    assert tab.shape == (n + 4, 4)
    assert not tab[n:].any(), "padding records must be zero"
    a = hs.g["pos"].astype(np.float32)[ids] - e
    assert np.array_equal(tab[:n, :3].view(np.uint32), a.view(np.uint32)), "a = mu - eye differs"
    # q on a sample of slots
    rng = np.random.default_rng(seed)
    pick = rng.choice(n, size=min(n, n_q), replace=False)
    checked = bounded = 0
    for j in pick:
        gid = int(ids[j])
        q = tab[j, 3]
        sh = shape(hs.g_inv[gid:gid + 1])
        if sh is None:  # no bound sphere: the record accepts every direction
            assert q == NO_BOUND, f"slot {j}: expected no bound"
            continue
        Smat, kap = sh
        kap = f16_up(kap)  # the stored kappa (the device's dilation of the ellipsoid)
        mu = hs.g_inv["pos"][gid]
        qe, a32 = q_ellipse(Smat, kap, mu, e)
        qs = q_sphere(Smat, kap, mu, e)
        na = float(np.linalg.norm(a32))
        tol = 1e-6 * na + 1e-7
        hi = max(qe, qs)
        if hi == -np.inf:
            assert q == NO_BOUND, f"slot {j}: no bound restated, device q {q}"
        else:
            assert float(q) <= hi + tol, f"slot {j}: q {q} above the bounds {qe}, {qs}"
        if qe > -np.inf:
            assert float(q) >= qe - tol, f"slot {j}: q {q} below the ellipsoid bound {qe}"
        checked += 1
        bounded += int(q > 0)
    return checked, bounded


@pytest.mark.parametrize("eye", [(0.0, 0.0, 1.6), (4.0, 3.0, 6.0), (0.2, -0.1, 0.3)])
def test_cone_table_anisotropic(dev, eye):
    checked, bounded = _check(dev, S.anisotropic(), eye, 1500, 3)
    assert checked > 1000
    if eye == (4.0, 3.0, 6.0):
        assert bounded == checked  # every Gaussian seen from far away has a cone


def test_cone_table_uniform_10k(dev):
    checked, bounded = _check(dev, S.uniform_10k(), (0.0, 0.0, 3.0), 1500, 5)
    assert checked > 1000 and bounded > 0.9 * checked


def test_cone_table_clustered_1m(dev):
    """9.4M slots: many chunks per wavefront of the copy, and its tail."""
    hs = S.clustered(1_000_000)
    aabb = np.asarray(hs.aabb, np.float64)
    c = 0.5 * (aabb[:3] + aabb[3:])
    eye = tuple((c + 25.0 * np.array([0.48, 0.36, 0.8])).astype(np.float32))
    checked, bounded = _check(dev, hs, eye, 2000, 7)
    assert checked > 1500 and bounded > 0.9 * checked
