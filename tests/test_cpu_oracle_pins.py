"""Independent pins of the oracle's intersection and SH evaluation.

Closed forms computed here in float64 from the mathematics (not from the
oracle's formulas) pin what the restatement computes; a literal float32
restatement of cudaRenderer.cu:517-549 pins how the reference rounds it.
Every case is chosen so that a plausible slip fails it:
  * a transposed rotation (R instead of R^T) or a swapped scale axis moves
    the rotated anisotropic Gaussian's t and grayDist far outside tolerance;
  * a wrong basis index, sign, channel stride or constant moves one of the
    15 single-coefficient cases;
  * double-precision constants or a product rounded once instead of twice
    breaks the bit-exact float32 comparison on a measurable share of
    directions.
"""
import importlib
import math

import numpy as np
import pytest

import oracle_lib as O

K = importlib.import_module("3dgrt_kdtree_amd")

# The reference's constants as decimal values (cudaRenderer.h:33-48).
C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435]


def basis_closed_form(d):
    """Real SH basis of degrees 0-3 in float64, with the reference's sign
    convention folded in (degree 1 enters as -y, +z, -x).  The polynomials
    are the textbook real spherical harmonics; the constants are
    sqrt((2l+1)/4pi * (l-m)!/(l+m)!) times the usual factors, checked below
    against their closed forms."""
    x, y, z = (float(v) for v in d)
    return np.array([
        C0,
        -C1 * y, C1 * z, -C1 * x,
        C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * z * z - x * x - y * y), C2[3] * x * z, C2[4] * (x * x - y * y),
        C3[0] * y * (3 * x * x - y * y), C3[1] * x * y * z, C3[2] * y * (4 * z * z - x * x - y * y),
        C3[3] * z * (2 * z * z - 3 * x * x - 3 * y * y), C3[4] * x * (4 * z * z - x * x - y * y),
        C3[5] * z * (x * x - y * y), C3[6] * x * (x * x - 3 * y * y)])


def basis_literal_f32(d):
    """cudaRenderer.cu:517-549 in numpy float32: f-suffixed constants, a * b * c
    evaluated as (a * b) * c, every product and sum rounded to float."""
    f = np.float32
    x, y, z = (f(v) for v in d)
    c1 = f(C1)
    c2 = [f(v) for v in C2]
    c3 = [f(v) for v in C3]
    xx, yy, zz = x * x, y * y, z * z
    xy, yz, xz = x * y, y * z, x * z
    return np.array([
        f(C0), c1 * y, c1 * z, c1 * x,
        c2[0] * xy, c2[1] * yz, c2[2] * (f(2) * zz - xx - yy), c2[3] * xz, c2[4] * (xx - yy),
        (c3[0] * y) * (f(3) * xx - yy), (c3[1] * xy) * z, (c3[2] * y) * (f(4) * zz - xx - yy),
        (c3[3] * z) * (f(2) * zz - f(3) * xx - f(3) * yy), (c3[4] * x) * (f(4) * zz - xx - yy),
        (c3[5] * z) * (xx - yy), (c3[6] * x) * (xx - f(3) * yy)], np.float32)


def basis_double_constants(d):
    """The misreading this suite guards against: double constants times float
    operands, rounded once to float."""
    f = np.float32
    x, y, z = (f(v) for v in d)
    xx, yy, zz = x * x, y * y, z * z
    xy, yz, xz = x * y, y * z, x * z
    D = np.float64
    return np.array([
        f(C0), f(C1 * D(y)), f(C1 * D(z)), f(C1 * D(x)),
        f(C2[0] * D(xy)), f(C2[1] * D(yz)), f(C2[2] * D(f(2) * zz - xx - yy)), f(C2[3] * D(xz)),
        f(C2[4] * D(xx - yy)),
        f(C3[0] * D(y) * D(f(3) * xx - yy)), f(C3[1] * D(xy) * D(z)), f(C3[2] * D(y) * D(f(4) * zz - xx - yy)),
        f(C3[3] * D(z) * D(f(2) * zz - f(3) * xx - f(3) * yy)), f(C3[4] * D(x) * D(f(4) * zz - xx - yy)),
        f(C3[5] * D(z) * D(xx - yy)), f(C3[6] * D(x) * D(xx - f(3) * yy))], np.float32)


def test_sh_constants_closed_form():
    # normalisation constants of the real SH (l, |m|): sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) times
    # sqrt(2) for m != 0, times the polynomial's own factor.
    pi = math.pi
    assert abs(C0 - math.sqrt(1 / (4 * pi))) < 1e-16
    assert abs(C1 - math.sqrt(3 / (4 * pi))) < 1e-16
    assert abs(C2[0] - 0.5 * math.sqrt(15 / pi)) < 1e-15 and C2[1] == C2[3] == -C2[0]
    assert abs(C2[2] - 0.25 * math.sqrt(5 / pi)) < 1e-15
    assert abs(C2[4] - 0.25 * math.sqrt(15 / pi)) < 1e-15
    assert abs(C3[0] + 0.25 * math.sqrt(35 / (2 * pi))) < 1e-15 and C3[6] == C3[0]
    assert abs(C3[1] - 0.5 * math.sqrt(105 / pi)) < 1e-14
    assert abs(C3[2] + 0.25 * math.sqrt(21 / (2 * pi))) < 1e-15 and C3[4] == C3[2]
    assert abs(C3[3] - 0.25 * math.sqrt(7 / pi)) < 1e-15
    assert abs(C3[5] - 0.25 * math.sqrt(105 / pi)) < 1e-14


def random_dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def test_sh_basis_closed_form_and_float_rounding():
    dirs = random_dirs(4000, 11)
    n_diff_double = 0
    for d in dirs:
        s = O.sh_basis(d)
        # signs of degree 1 are applied by eval_sh; compare magnitudes there
        ref = basis_closed_form(d.astype(np.float64))
        ref[1], ref[3] = -ref[1], -ref[3]
        assert np.allclose(s.astype(np.float64), ref, rtol=3e-6, atol=3e-7), (d, s, ref)
        lit = basis_literal_f32(d)
        assert np.array_equal(s.view(np.uint32), lit.view(np.uint32)), (d, s, lit)
        n_diff_double += int(not np.array_equal(lit, basis_double_constants(d)))
    # the double-constant reading differs from the reference's float reading
    # on a large share of directions, so the bit-exact check above catches it
    assert n_diff_double > 0.1 * len(dirs), n_diff_double


def sh_record(f_dc=(0.0, 0.0, 0.0), rest=None):
    g = np.zeros(1, dtype=K.GAUSSIAN_DTYPE)
    g["f_dc"] = f_dc
    if rest is not None:
        g["f_rest"] = rest
    g["opacity"] = 0.5
    g["scale"] = 10.0
    g["rot"] = (1, 0, 0, 0)
    return g


DIRS = [(1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1), (1, -2, 3), (-3, 1, 2), (0.2, -0.7, -0.4)]


@pytest.mark.parametrize("k", range(15))
def test_sh_each_coefficient_closed_form(k):
    """Only f_rest coefficient k+1 is set, with distinct values per channel
    (channel-major: f_rest[k], f_rest[15+k], f_rest[30+k]).  The colour is
    max(0.5 + sign * basis_k(d) * f, 0) per channel."""
    vals = (0.8, -0.45, 0.3)
    rest = np.zeros(45, np.float32)
    rest[k], rest[15 + k], rest[30 + k] = vals
    g = sh_record(rest=rest)
    for dd in DIRS:
        d = np.asarray(dd, np.float64)
        d /= np.linalg.norm(d)
        d32 = d.astype(np.float32)
        b = basis_closed_form(d32.astype(np.float64))[k + 1]
        exp = np.maximum(0.5 + b * np.asarray(vals), 0.0)
        got = O.eval_sh(d32, g)
        assert np.allclose(got, exp, atol=2e-6), (k, dd, got, exp)
        got64 = O.eval_sh(d32.astype(np.float64), g, f64=True)
        assert np.allclose(got64, exp, atol=1e-12), (k, dd, got64, exp)


def test_sh_full_record_float_order():
    """All 16 coefficients set: eval_sh equals the reference's float32
    accumulation rad = SH_C0 c0 - s1 c1 + s2 c2 - s3 c3 + s4 c4 + ... + 0.5,
    clamped below at 0, bit for bit; and the FP64 evaluation within float
    rounding."""
    rng = np.random.default_rng(5)
    f = np.float32
    for d in random_dirs(500, 12):
        dc = rng.normal(size=3).astype(np.float32)
        rest = (0.3 * rng.normal(size=45)).astype(np.float32)
        g = sh_record(dc, rest)
        s = basis_literal_f32(d)
        exp = np.zeros(3, np.float32)
        for ch in range(3):
            c = [dc[ch]] + [rest[15 * ch + i] for i in range(15)]
            r = c[0] * s[0]
            r = r - c[1] * s[1]
            r = r + c[2] * s[2]
            r = r - c[3] * s[3]
            for i in range(4, 16):
                r = r + c[i] * s[i]
            r = r + f(0.5)
            exp[ch] = max(r, f(0))
        got = O.eval_sh(d, g)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (d, got, exp)
        b = basis_closed_form(d.astype(np.float64))
        c64 = np.stack([np.concatenate([[dc[ch]], rest[15 * ch:15 * ch + 15]]) for ch in range(3)]).astype(np.float64)
        ref = np.maximum(c64 @ b + 0.5, 0.0)
        assert np.allclose(got, ref, atol=2e-5), (got, ref)


def rot_z(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]])


def mahalanobis_closest(o, d, mu, R, s):
    """Closest approach of the ray o + t d to the Gaussian N(mu, R diag(s^2) R^T)
    in the Mahalanobis metric: t* = d^T P (mu - o) / d^T P d with P the
    inverse covariance, and the squared distance there.  For |d| = 1, t* is
    the world distance the reference reports as t."""
    P = R @ np.diag(1.0 / np.asarray(s) ** 2) @ R.T
    t = d @ P @ (mu - o) / (d @ P @ d)
    e = o + t * d - mu
    return t, e @ P @ e


def gauss_record(mu, scale, quat_wxyz, opacity=0.9):
    g = np.zeros(1, dtype=K.GAUSSIAN_DTYPE)
    g["pos"] = mu
    g["scale"] = 1.0 / np.asarray(scale, np.float32)   # the device reads inverted scales
    g["rot"] = quat_wxyz
    g["opacity"] = opacity
    return g


def quat_to_matrix(q):
    """Rotation matrix of a unit quaternion (w, x, y, z), Hamilton convention."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_rotated_anisotropic_gaussian_30deg_about_z():
    """Scales (0.3, 0.05, 0.05), long axis rotated 30 degrees about z.  The
    rotation matrix is built from the angle, not from the quaternion."""
    mu = np.array([0.2, -0.1, 0.05])
    s = (0.3, 0.05, 0.05)
    h = math.radians(15.0)
    g = gauss_record(mu, s, (math.cos(h), 0.0, 0.0, math.sin(h)))
    R = rot_z(30.0)
    rays = [((0.35, 0.02, -3.0), (0.0, 0.0, 1.0)),        # down the z axis, off-centre along the long axis
            ((0.45, 0.1, 2.5), (0.0, 0.0, -1.0)),
            ((-2.0, 0.6, 0.05), (1.0, -0.3, 0.0)),        # oblique, in the xy plane
            ((0.2, -3.0, 0.3), (0.05, 1.0, -0.1)),
            ((1.5, 1.2, 1.0), (-0.6, -0.5, -0.4))]
    n_transposed_differs = 0
    for o, d in rays:
        o = np.asarray(o, np.float64)
        d = np.asarray(d, np.float64)
        d /= np.linalg.norm(d)
        d32, o32 = d.astype(np.float32), o.astype(np.float32)
        t_ref, gd_ref = mahalanobis_closest(o32.astype(np.float64), d32.astype(np.float64), mu.astype(np.float32)
                                            .astype(np.float64), R, s)
        acc64, t64, gd64 = O.ray_gaussian(o32.astype(np.float64), d32.astype(np.float64), g, f64=True)
        # the record holds f32 inverse scales and quaternion: compare within their rounding
        assert abs(t64 - t_ref) <= 1e-6 * max(1.0, abs(t_ref)), (o, d, t64, t_ref)
        assert abs(gd64 - gd_ref) <= 1e-5 * max(1.0, gd_ref), (o, d, gd64, gd_ref)
        acc32, t32, gd32 = O.ray_gaussian(o32, d32, g)
        assert abs(t32 - t_ref) <= 2e-5 * max(1.0, abs(t_ref)), (o, d, t32, t_ref)
        assert abs(gd32 - gd_ref) <= 2e-4 * max(1.0, gd_ref), (o, d, gd32, gd_ref)
        assert acc32 == (gd_ref < 8 and t_ref >= 1e-2), (o, d, acc32, gd_ref, t_ref)
        t_tr, gd_tr = mahalanobis_closest(o, d, mu, R.T, s)
        n_transposed_differs += int(abs(gd_tr - gd_ref) > 0.05 * max(1.0, gd_ref))
        # swapping the long axis onto y changes the answer too
        t_sw, gd_sw = mahalanobis_closest(o, d, mu, R, (s[1], s[0], s[2]))
        assert abs(gd_sw - gd_ref) > 0.05 * max(1.0, gd_ref) or abs(t_sw - t_ref) > 1e-3
    assert n_transposed_differs >= 3


def test_random_anisotropic_gaussians_fp64():
    """Random unit quaternions and anisotropic scales: the FP64 oracle equals
    the Mahalanobis closest approach to 1e-9 (relative), the FP32 oracle to
    float rounding."""
    rng = np.random.default_rng(21)
    for _ in range(300):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        q32 = q.astype(np.float32)
        s = np.exp(rng.normal(np.log(0.1), 0.8, 3))
        mu = rng.uniform(-1, 1, 3)
        g = gauss_record(mu, s, q32)
        # the quantities the record actually holds
        s_rec = 1.0 / g["scale"][0].astype(np.float64)
        mu_rec = g["pos"][0].astype(np.float64)
        R = quat_to_matrix(q32.astype(np.float64))
        o = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        t_ref, gd_ref = mahalanobis_closest(o, d, mu_rec, R, s_rec)
        acc64, t64, gd64 = O.ray_gaussian(o, d, g, t_near=-1e30, t_far=1e30, f64=True)
        # a float quaternion is not exactly unit: R is then a scaled rotation;
        # compare against the orthogonal closed form within that defect
        defect = abs(float(np.sum(q32.astype(np.float64) ** 2)) - 1.0)
        tol = 1e-9 + 8 * defect
        assert abs(t64 - t_ref) <= tol * max(1.0, abs(t_ref)) * 10, (t64, t_ref, defect)
        assert abs(gd64 - gd_ref) <= tol * max(1.0, gd_ref) * 10, (gd64, gd_ref, defect)
        if gd_ref < 50:
            acc32, t32, gd32 = O.ray_gaussian(o.astype(np.float32), d.astype(np.float32), g, t_near=-1e30,
                                              t_far=1e30)
            assert abs(t32 - t_ref) <= 1e-4 * max(1.0, abs(t_ref)), (t32, t_ref)
            assert abs(gd32 - gd_ref) <= 1e-3 * max(1.0, gd_ref), (gd32, gd_ref)
