"""Pins the oracle's level-set colliders (ApplyGridBoundaryOnZSGrid's response, Boundary.cu:55-66,
with zpc Collider's frame and object velocity) by closed forms:
  sticky      inside -> v = v_obj = translation_rate + omega x (x - b) + scale_rate / scale (x - b);
  slip        the normal component of v - v_obj is removed, the tangential part kept;
  separate    removed only when v - v_obj points into the object;
  outside     (phi >= 0) v unchanged;
  frames      yaw/pitch/roll rotate the normal as R = Ry Rx Rz; scale scales distances;
  shapes      sphere, box (face normals, outside distance), half-space; a sampled SDF of a linear
              function equals the half-space exactly (trilinear interpolation reproduces linear
              functions) and reports no contact outside its sample box.
"""
import numpy as np
import pytest

from oracle import oracle as O

HALF, SPHERE, BOX, GRID = 0, 1, 2, 3
STICKY, SLIP, SEP = 0, 1, 2


def rx(deg):
    a = np.radians(deg)
    return np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])


def ry(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])


def rz(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])


def test_sticky_object_velocity(oracle):
    col = dict(type=STICKY, shape=SPHERE, params=(0, 0, 0, 0.1), translation=(0.5, 0.5, 0.5),
               translation_rate=(0.3, -0.2, 0.1), angular_velocity=(0.0, 2.0, 0.0), scale=2.0, scale_rate=0.5)
    x = np.array([0.6, 0.55, 0.45])  # |x - b| = 0.122 < 0.2 = scale * r
    d = x - np.array([0.5, 0.5, 0.5])
    vo = np.array([0.3, -0.2, 0.1]) + np.cross([0.0, 2.0, 0.0], d) + 0.5 / 2.0 * d
    v, inside = O.resolve_collision(col, x, [5.0, 6.0, 7.0])
    assert inside and np.allclose(v, vo, atol=1e-6)
    v, inside = O.resolve_collision(col, [0.75, 0.5, 0.5], [5.0, 6.0, 7.0])  # 0.25 > 0.2: outside
    assert not inside and np.array_equal(v, [5.0, 6.0, 7.0])


def test_slip_and_separate_halfspace(oracle):
    base = dict(shape=HALF, params=(0, 1, 0, 0.1))
    x = [0.3, 0.05, 0.7]
    v, inside = O.resolve_collision(dict(base, type=SLIP), x, [1.0, -2.0, 0.5])
    assert inside and np.allclose(v, [1.0, 0.0, 0.5])
    v, _ = O.resolve_collision(dict(base, type=SLIP), x, [1.0, 3.0, 0.5])
    assert np.allclose(v, [1.0, 0.0, 0.5])
    v, _ = O.resolve_collision(dict(base, type=SEP), x, [1.0, -2.0, 0.5])
    assert np.allclose(v, [1.0, 0.0, 0.5])
    v, _ = O.resolve_collision(dict(base, type=SEP), x, [1.0, 3.0, 0.5])
    assert np.allclose(v, [1.0, 3.0, 0.5])
    v, inside = O.resolve_collision(dict(base, type=SLIP), [0.3, 0.15, 0.7], [1.0, -2.0, 0.5])
    assert not inside and np.array_equal(v, [1.0, -2.0, 0.5])
    # a moving plane: only the normal part of the relative velocity goes
    v, _ = O.resolve_collision(dict(base, type=SLIP, translation_rate=(0.0, 0.4, 0.0)), x, [1.0, -2.0, 0.5])
    assert np.allclose(v, [1.0, 0.4, 0.5])


def test_rotated_frames(oracle):
    ypr = (25.0, 30.0, -40.0)
    R = ry(ypr[0]) @ rx(ypr[1]) @ rz(ypr[2])
    n = R @ np.array([0.0, 1.0, 0.0])
    b = np.array([0.4, 0.5, 0.6])
    col = dict(type=SLIP, shape=HALF, params=(0, 1, 0, 0.0), translation=b, ypr_degrees=ypr)
    x = b - 0.03 * n + 0.02 * np.cross(n, [1.0, 0.0, 0.0])
    vin = np.array([0.7, -1.1, 0.4])
    v, inside = O.resolve_collision(col, x, vin)
    assert inside and np.allclose(v, vin - np.dot(vin, n) * n, atol=1e-6)
    v, inside = O.resolve_collision(col, b + 0.03 * n, vin)
    assert not inside
    # a rotated, scaled box: the nearest face's normal
    col = dict(type=SLIP, shape=BOX, params=(-0.1, -0.2, -0.3, 0.1, 0.2, 0.3), translation=b, ypr_degrees=ypr,
               scale=0.5)
    ex = R @ np.array([1.0, 0.0, 0.0])
    x = b + 0.5 * 0.09 * ex  # 0.01 inside the +x face (scaled by 0.5)
    v, inside = O.resolve_collision(col, x, vin)
    assert inside and np.allclose(v, vin - np.dot(vin, ex) * ex, atol=1e-6)
    v, inside = O.resolve_collision(col, b + 0.5 * 0.11 * ex, vin)
    assert not inside


def test_sampled_plane_equals_halfspace(oracle):
    n = np.array([0.3, 0.8, -0.2])
    n /= np.linalg.norm(n)
    h, org = 0.05, np.array([0.1, 0.1, 0.1])
    idx = np.stack(np.meshgrid(*[np.arange(12)] * 3, indexing="ij"), axis=-1)
    pts = org + h * idx
    sdf = (pts @ n - 0.4).astype(np.float32)
    plane = dict(type=SEP, shape=HALF, params=(*n, 0.4))
    grid = dict(type=SEP, shape=GRID, sdf=sdf, origin=org, spacing=h)
    rng = np.random.default_rng(8)
    hits = 0
    for _ in range(200):
        x = org + rng.uniform(0, 11 * h, size=3)
        vin = rng.normal(size=3)
        a, ia = O.resolve_collision(plane, x, vin)
        b, ib = O.resolve_collision(grid, x, vin)
        if abs(x @ n - 0.4) < 1e-5:
            continue
        assert ia == ib
        assert np.allclose(a, b, atol=1e-5)
        hits += ia
    assert hits > 20
    v, inside = O.resolve_collision(grid, org - 0.01, [1.0, 2.0, 3.0])  # outside the samples
    assert not inside
