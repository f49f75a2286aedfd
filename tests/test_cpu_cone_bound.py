"""Pins the view pre-test's cone threshold against the FP32 oracle's exact test.

A view's cone record for a Gaussian is (a, q) with a = mu - e in f32 and q a
lower bound of the f32 dot product a.d over every ray direction d from the
eye e that the exact test (cudaRenderer.cu:1069-1135, restated in the oracle)
accepts; k_gaussian_cone forms q as the larger of a bound-sphere bound and a
projected-ellipsoid bound.  This file restates the projected-ellipsoid bound
in float64 from its geometric derivation (the f32 rotation of the device,
S = 8 (A^T A)^-1 with A = diag(1/scale) M stored in f32, the same margins)
and checks it on anisotropic Gaussians (axis ratios up to 100) seen from
near and far eyes, with rays aimed at points on and just outside the
acceptance ellipsoid, so most accepted rays graze its boundary:

* every ray the FP32 exact test accepts has a.d >= q + 4 eps |a| in exact
  arithmetic on the f32 inputs, so the device's three-rounding f32 dot
  product passes cone_pass;
* the same check with the projected extent taken from the smallest instead
  of the largest in-plane eigenvalue, or with the angle bounded at the
  centre's distance instead of the ellipsoid's near face, fails: the rays
  reach the boundary where the bound is tight;
* the ellipsoid bound is tighter than the bound-sphere one on average.
"""
import math

import numpy as np

import oracle_lib as O
from test_cpu_oracle_pins import gauss_record

EPS = 2.0 ** -24


def mt_f32(q):
    """quaternionWXYZToMatrixTranspose (cudaRenderer.cu:808-828) in f32."""
    f = np.float32
    qr, qx, qy, qz = (f(v) for v in q)
    xx, yy, zz = qx * qx, qy * qy, qz * qz
    xy, xz, yz = qx * qy, qx * qz, qy * qz
    rx, ry, rz = qr * qx, qr * qy, qr * qz
    two, one = f(2), f(1)
    return np.array([[one - two * (yy + zz), two * (xy + rz), two * (xz - ry)],
                     [two * (xy - rz), one - two * (xx + zz), two * (yz + rx)],
                     [two * (xz + ry), two * (yz - rx), one - two * (xx + yy)]], np.float32)


def shape(rec):
    """(S in f32 as the device stores it, kappa) or None for r = inf."""
    g = rec["scale"][0].astype(np.float64)  # inverse scales
    M = mt_f32(rec["rot"][0]).astype(np.float64)
    if np.abs(M.T @ M - np.eye(3)).max() > 4e-6:
        return None
    A = np.diag(g) @ M
    P = A.T @ A
    ev = np.linalg.eigvalsh(P)
    kap = 8e-6 * (4.0 + math.sqrt(ev[-1] / ev[0]))
    S = 8.0 * np.linalg.inv(P)
    return S.astype(np.float32).astype(np.float64), kap


def q_ellipse(S, kap, mu, e, lam="max", slab=True):
    """The projected-ellipsoid threshold of k_gaussian_cone (float64)."""
    a32 = (mu.astype(np.float32) - e.astype(np.float32)).astype(np.float64)
    l1 = np.abs(a32).sum()
    m = mu.astype(np.float64) - e.astype(np.float64)
    nm = np.linalg.norm(m)
    v = m / nm
    tr = np.trace(S)
    Pm = np.eye(3) - np.outer(v, v)
    ev = np.linalg.eigvalsh(Pm @ S @ Pm)
    lp = (ev[-1] if lam == "max" else ev[1]) + 2e-7 * tr  # ev[0] ~ 0 is the v direction
    vsv = max(v @ S @ v, 0.0)
    dil = kap * l1
    Rc = 1.001 * math.sqrt(lp) + dil
    h = 1.001 * math.sqrt(vsv + 2e-7 * tr) + dil
    z0 = nm - h if slab else nm
    if z0 <= 0:
        return -np.inf, a32
    C = z0 / math.sqrt(z0 * z0 + Rc * Rc) - 12 * EPS
    return (nm * C if C > 0 else -np.inf), a32


def q_sphere(S, kap, mu, e):
    """The bound-sphere threshold (cone record of the sphere test's R)."""
    r = math.sqrt(np.linalg.eigvalsh(S)[-1]) * 1.001
    a32 = (mu.astype(np.float32) - e.astype(np.float32)).astype(np.float64)
    R = r + kap * np.abs(a32).sum()
    a2 = a32 @ a32
    na = math.sqrt(a2)
    Rp = R * (1 + 4 * EPS) + 10 * EPS * na
    Q2 = a2 * (1 - 16 * EPS) - Rp * Rp
    lim = R + 4 * EPS * na
    return math.sqrt(Q2) - 4 * EPS * na if Q2 > 0 and Q2 >= lim * lim else -np.inf


def f32_dir(e, p):
    """A view ray's f32 direction towards p (make_ray: w = p - e, d = w / |w|)."""
    w = (p.astype(np.float32) - e.astype(np.float32)).astype(np.float32)
    inv = np.float32(1.0) / np.sqrt(np.float32(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]))
    return (w * inv).astype(np.float32)


def _cases(n_gauss, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n_gauss):
        s = np.exp(rng.normal(math.log(0.05), 0.9, 3))
        s[rng.integers(0, 3)] *= rng.choice([1.0, 10.0, 100.0])  # needles and pancakes
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        mu = rng.uniform(-2, 2, 3)
        rec = gauss_record(mu, s, q.astype(np.float32))
        sh = shape(rec)
        if sh is None:
            continue
        S, kap = sh
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = math.exp(rng.uniform(math.log(1.5), math.log(30.0))) * max(1.0, 3 * math.sqrt(np.trace(S)))
        e = (rec["pos"][0].astype(np.float64) + dist * d).astype(np.float32)
        yield rec, S, kap, rng, e


def _boundary_rays(rec, S, rng, e, n):
    """Rays from e towards points mu + (1 + delta) sqrt(8) A^-1 u, |u| = 1,
    delta in [-2e-3, 2e-3]: on and just outside the acceptance ellipsoid."""
    mu = rec["pos"][0].astype(np.float64)
    L = np.linalg.cholesky(S)  # S = L L^T: mu + L u lies on the ellipsoid for |u| = 1
    for _ in range(n):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        p = mu + (1.0 + rng.uniform(-2e-3, 2e-3)) * (L @ u)
        yield f32_dir(e, p)


def _check(lam="max", slab=True, n_gauss=300, rays=60, seed=31):
    accepted = violations = 0
    tighter = []
    for rec, S, kap, rng, e in _cases(n_gauss, seed):
        mu = rec["pos"][0]
        q, a32 = q_ellipse(S, kap, mu, e, lam, slab)
        qs = q_sphere(S, kap, mu, e)
        na = np.linalg.norm(a32)
        if np.isfinite(q) and np.isfinite(qs):
            tighter.append((na - q) / (na - qs))
        for d in _boundary_rays(rec, S, rng, e, rays):
            acc, t, gd = O.ray_gaussian(e, d, rec)
            if not acc:
                continue
            accepted += 1
            dot = a32 @ d.astype(np.float64)
            violations += int(np.isfinite(q) and dot < q + 4 * EPS * na)
    return accepted, violations, np.array(tighter)


def test_ellipse_cone_bound_is_conservative():
    accepted, violations, tighter = _check()
    print(f"accepted boundary rays {accepted}, violations {violations}, "
          f"median (|a| - q) ratio ellipse / sphere {np.median(tighter):.3f}")
    assert accepted > 3000
    assert violations == 0
    # the cone depth |a| - q shrinks against the bound sphere's
    assert np.median(tighter) < 0.95


def test_planted_mistakes_fail():
    # the projected extent from the smallest in-plane eigenvalue
    _, v_min, _ = _check(lam="min", n_gauss=150)
    assert v_min > 0
    # the angle bounded at the centre's distance instead of the near face
    # of the slab the ellipsoid occupies along the view direction
    _, v_noslab, _ = _check(slab=False, n_gauss=300)
    assert v_noslab > 0
