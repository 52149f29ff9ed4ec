"""Pins the environment-map oracle (oracle/envmap.py): lat-long coordinates of the axes, the
rotation convention against the elementary rotations, and bilinear filtering edge rules."""
import numpy as np
import pytest

from oracle import envmap as O


def test_latlong_axes():
    uv = O.latlong_uv(np.array([[0, 0, -1], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [-1, 0, 0]], np.float64))
    # -Z at the centre, +X at u = 3/4, +Y on the top row, -Y on the bottom row, +Z at the seam.
    np.testing.assert_allclose(uv[0], [0.5, 0.5], atol=1e-15)
    np.testing.assert_allclose(uv[1], [0.75, 0.5], atol=1e-15)
    assert uv[2, 1] == 0.0 and uv[3, 1] == pytest.approx(1.0)
    assert uv[4, 0] in (0.0, 1.0) and uv[5, 0] == pytest.approx(0.25)


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_rotation_convention():
    # matrixFromRotationXYZ(x, y, z) = Rx(x) Ry(y) Rz(z) for column vectors (mul(M, v)).
    ang = np.array([20.0, -35.0, 50.0])
    R = O.rotation_xyz(ang)
    a = np.radians(ang)
    np.testing.assert_allclose(R, _rx(a[0]) @ _ry(a[1]) @ _rz(a[2]), atol=1e-14)
    # A quarter turn about Y takes -Z (the map centre) to -X.
    np.testing.assert_allclose(O.rotation_xyz([0, 90, 0]) @ [0, 0, -1], [-1, 0, 0], atol=1e-15)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)


def test_bilinear_wrap_and_clamp():
    tex = np.zeros((2, 4, 3))
    tex[:, 0] = 1.0  # column 0 bright
    # u = 0: halfway between column 3 and column 0 (wrap).
    assert O.bilinear(tex, np.array([[0.0, 0.5]]))[0, 0] == pytest.approx(0.5)
    # v = 0 clamps to row 0; texel centres return texel values.
    assert O.bilinear(tex, np.array([[0.125, 0.0]]))[0, 0] == pytest.approx(1.0)
    tex2 = np.arange(2 * 4 * 3, dtype=np.float64).reshape(2, 4, 3)
    np.testing.assert_allclose(O.bilinear(tex2, np.array([[0.375, 0.75]]))[0], tex2[1, 1])


def test_constant_map_and_scaling():
    tex = np.full((8, 16, 3), 2.0)
    d = np.random.default_rng(0).standard_normal((100, 3))
    out = O.env_eval(tex, d, intensity=3.0, tint=(1.0, 0.5, 0.25), rotation=(10, 20, 30))
    np.testing.assert_allclose(out, np.broadcast_to([6.0, 3.0, 1.5], out.shape), rtol=1e-12)
