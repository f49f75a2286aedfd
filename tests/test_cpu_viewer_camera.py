"""Viewer camera controls (kdgrt_camera_move / orbit / drag) on the CPU.

Pinned by closed forms and invariants rather than by restating the code:
  * move (idle(), Kd-treeConverterMain.cpp:3463-3500): each key alone moves
    the eye by speed*dt along its axis with the reference's sign (w forward
    = -n, s +n, a -u, d +u, q +v, e -v); no key leaves the camera bit-equal.
  * orbit (updateOrbitCamera, :438-480): against a float64 rotation matrix
    (Rodrigues) for a unit axis; right-handed (x about z by +90 deg -> y);
    distance to the axis and the axial coordinate are kept; 360 steps of
    1 deg come back to the start; a non-unit axis is used as given (the
    hand-computed case (1,0,0) about (0,2,0) by 90 deg -> (0,0,-2)).
  * drag (mousemove, :1193-1296): yaw by delx*rot_speed degrees about world
    y and pitch about u, in closed form from a camera looking down -z; a
    positive pitch is blocked while n.y > 0.98 and a negative one while
    n.y < -0.98 (both sides checked); zoom changes fovy by delx/4 within
    [0.5, 100].
  * every rotation leaves u, v, n orthonormal and right-handed (u x v = n),
    and a zero-length n is zeroed as fMyVecNormalize does.
"""
import importlib
import math

import numpy as np
import pytest

K = importlib.import_module("3dgrt_kdtree_amd")


def cam_down_z(pos=(0.0, 0.0, 5.0)):
    c = K.Camera()
    for k in range(3):
        c.pos[k] = pos[k]
        c.u[k] = (1.0, 0.0, 0.0)[k]
        c.v[k] = (0.0, 1.0, 0.0)[k]
        c.n[k] = (0.0, 0.0, 1.0)[k]
    c.fovy_deg, c.aspect, c.near_c = 60.0, 1.0, 0.005
    return c


def vec(a):
    return np.array([a[0], a[1], a[2]], np.float64)


def frame(c):
    return vec(c.pos), vec(c.u), vec(c.v), vec(c.n)


def assert_orthonormal(c, tol=2e-6):
    _, u, v, n = frame(c)
    for a in (u, v, n):
        assert abs(np.linalg.norm(a) - 1.0) < tol
    assert abs(u @ v) < tol and abs(v @ n) < tol and abs(u @ n) < tol
    assert np.allclose(np.cross(u, v), n, atol=4 * tol)


def rodrigues(axis, theta):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) * math.cos(theta) + math.sin(theta) * k + (1 - math.cos(theta)) * np.outer(a, a)


@pytest.mark.parametrize("key,axis,sign", [(K.KEY_W, "n", -1), (K.KEY_S, "n", 1), (K.KEY_A, "u", -1),
                                           (K.KEY_D, "u", 1), (K.KEY_Q, "v", 1), (K.KEY_E, "v", -1)])
def test_move_each_key(key, axis, sign):
    c = K.look_at((1.0, 2.0, 3.0), (0.2, -0.1, 0.4), 60.0, 1.5)
    p0, u, v, n = frame(c)
    ax = {"u": u, "v": v, "n": n}[axis]
    speed, dt = 0.5, 1.0 / 60.0
    assert K.camera_move(c, key, speed, dt)
    d = vec(c.pos) - p0
    assert abs(np.linalg.norm(d) - speed * dt) < 1e-6
    assert np.allclose(d / np.linalg.norm(d), sign * ax, atol=1e-4)
    assert all(np.array_equal(a, b) for a, b in zip(frame(c)[1:], (u, v, n)))   # axes untouched


def test_move_no_key_and_opposite_keys():
    c = K.look_at((1.0, 2.0, 3.0), (0.0, 0.0, 0.0))
    before = bytes(c)
    assert not K.camera_move(c, 0, 0.5, 0.1)
    assert bytes(c) == before
    assert K.camera_move(c, K.KEY_W | K.KEY_S | K.KEY_A | K.KEY_D, 0.5, 0.1)
    assert np.allclose(vec(c.pos), (1.0, 2.0, 3.0), atol=1e-6)


def test_orbit_closed_form_and_invariants():
    rng = np.random.default_rng(7)
    centre = np.array([0.3, -0.2, 0.5])
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    c = K.look_at((2.0, 1.0, -1.5), (0.0, 0.0, 0.0), 50.0, 1.0)
    p0, u0, v0, n0 = frame(c)
    theta = 0.37
    K.camera_orbit(c, theta, tuple(centre), tuple(axis))
    R = rodrigues(axis, theta)
    p1, u1, v1, n1 = frame(c)
    assert np.allclose(p1, centre + R @ (p0 - centre), atol=2e-6)
    for a, b in ((u1, u0), (v1, v0), (n1, n0)):
        assert np.allclose(a, R @ b, atol=2e-6)
    # distance to the axis and the axial coordinate are kept
    r0, r1 = p0 - centre, p1 - centre
    assert abs(r0 @ axis - r1 @ axis) < 2e-6
    assert abs(np.linalg.norm(np.cross(r0, axis)) - np.linalg.norm(np.cross(r1, axis))) < 2e-6
    assert_orthonormal(c)


def test_orbit_right_handed_and_full_turn():
    c = cam_down_z(pos=(1.0, 0.0, 0.0))
    K.camera_orbit(c, math.pi / 2, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    assert np.allclose(vec(c.pos), (0.0, 1.0, 0.0), atol=1e-6)   # x -> y about +z
    assert np.allclose(vec(c.u), (0.0, 1.0, 0.0), atol=1e-6)
    c = K.look_at((3.0, 1.0, 2.0), (0.0, 0.0, 0.0))
    start = frame(c)
    for _ in range(360):
        K.camera_orbit(c, math.radians(1.0), K.ORBIT_CENTRE, K.ORBIT_AXIS)
    for a, b in zip(frame(c), start):
        assert np.allclose(a, b, atol=5e-4)
    assert_orthonormal(c)


def test_orbit_axis_used_as_given():
    # rotateVectorAroundArbitraryAxis does not normalise: with axis (0, 2, 0)
    # and 90 deg, (1, 0, 0) -> (1, 0, 0) cos + (a x p) sin + a (a.p)(1 - cos)
    # = (0, 0, -2) up to cos(pi/2) rounding.
    c = cam_down_z(pos=(1.0, 0.0, 0.0))
    K.camera_orbit(c, math.pi / 2, (0.0, 0.0, 0.0), (0.0, 2.0, 0.0))
    assert np.allclose(vec(c.pos), (0.0, 0.0, -2.0), atol=1e-6)
    assert_orthonormal(c)                                        # re-orthonormalised after the scaling


def test_drag_yaw_pitch_closed_form():
    c = cam_down_z()
    K.camera_drag(c, 100, 0, 0.1)                                # yaw 10 deg about +y
    s, co = math.sin(math.radians(10)), math.cos(math.radians(10))
    assert np.allclose(vec(c.n), (s, 0.0, co), atol=1e-6)
    assert np.allclose(vec(c.u), (co, 0.0, -s), atol=1e-6)
    assert np.allclose(vec(c.v), (0.0, 1.0, 0.0), atol=1e-6)
    assert np.array_equal(vec(c.pos), (0.0, 0.0, 5.0))           # orientation only
    c = cam_down_z()
    K.camera_drag(c, 0, 50, 0.1)                                 # pitch 5 deg about u = +x
    s, co = math.sin(math.radians(5)), math.cos(math.radians(5))
    assert np.allclose(vec(c.n), (0.0, -s, co), atol=1e-6)
    assert np.allclose(vec(c.v), (0.0, co, s), atol=1e-6)
    assert_orthonormal(c)


def test_drag_pitch_clamp():
    # the reference blocks a positive pitch while n.y > 0.98 and a negative
    # one while n.y < -0.98; a positive pitch lowers n.y (n points backwards,
    # so the rule blocks the turn away from the pole, not towards it)
    c = cam_down_z()
    K.camera_drag(c, 0, -800, 0.1)                               # -80 deg: n.y = sin 80 deg
    assert abs(vec(c.n)[1] - math.sin(math.radians(80))) < 1e-6
    n_before = vec(c.n)
    K.camera_drag(c, 0, 10, 0.1)                                 # positive pitch at n.y > 0.98: blocked
    assert np.allclose(vec(c.n), n_before, atol=1e-6)
    K.camera_drag(c, 0, -10, 0.1)                                # negative pitch: turns on, to 81 deg
    assert abs(vec(c.n)[1] - math.sin(math.radians(81))) < 1e-5
    c = cam_down_z()
    K.camera_drag(c, 0, 800, 0.1)                                # +80 deg: n.y = -sin 80 deg
    n_before = vec(c.n)
    K.camera_drag(c, 0, -10, 0.1)                                # negative pitch at n.y < -0.98: blocked
    assert np.allclose(vec(c.n), n_before, atol=1e-6)
    K.camera_drag(c, 0, 10, 0.1)
    assert abs(vec(c.n)[1] + math.sin(math.radians(81))) < 1e-5
    assert_orthonormal(c)


def test_drag_zoom():
    c = cam_down_z()
    axes = bytes(c)[:48]
    K.camera_drag(c, 10, 99, zoom=True)
    assert c.fovy_deg == 62.5 and bytes(c)[:48] == axes
    K.camera_drag(c, 1000, 0, zoom=True)
    assert c.fovy_deg == 100.0
    K.camera_drag(c, -1000, 0, zoom=True)
    assert c.fovy_deg == 0.5


def test_reorthonormalise_skewed_axes_and_zero_n():
    c = cam_down_z()
    for k, val in enumerate((0.1, 0.05, 1.2)):
        c.n[k] = val
    c.u[1] = 0.3                                                 # u not orthogonal, not unit
    n_dir = vec(c.n) / np.linalg.norm(vec(c.n))
    K.camera_drag(c, 0, 0, 0.1)                                  # no rotation: only re-orthonormalised
    assert np.allclose(vec(c.n), n_dir, atol=1e-6)
    assert_orthonormal(c)
    z = cam_down_z()
    for k in range(3):
        z.n[k] = 0.0
    K.camera_orbit(z, 0.1, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert np.all(vec(z.n) == 0) and np.all(vec(z.v) == 0) and np.all(vec(z.u) == 0)
