"""Pins the oracle's sparse-leaf SDF collider shape (the SparseGrid that ToZSLevelSet builds from a
VDB FloatGrid, projects/CuLagrange/mpm/Generation.cpp:1536-1547; zpc's SparseGrid and OpenVDB are
absent from the reference tree).  The leaf layout is OpenVDB's LeafNode<float, 3>: 8^3 samples,
sample (x, y, z) at offset (x & 7) * 64 + (y & 7) * 8 + (z & 7).  Checks:
  - a linear function stored in leaves is reproduced exactly by the trilinear interpolant (value
    and gradient): a transposed offset order, a wrong leaf lookup or a wrong origin breaks it;
  - leaves that tile a box equal the dense SDF grid over the same samples, bit for bit;
  - leaf order does not matter (shuffling leaves with their samples);
  - negative leaf origins (floor division of negative sample indices);
  - samples outside every leaf read the background: all eight corners outside -> phi = background
    and a zero gradient; a mixed cell matches brute force over a dict of samples.
"""
import numpy as np

from oracle import oracle as O

GRID, LEAVES = 3, 4
N = np.array([0.3, -0.8, 0.5])
N = N / np.linalg.norm(N)


def leaves_of(dense, leaf_lo, nleaf):
    """Cut a dense (8 nx, 8 ny, 8 nz) array whose sample (0, 0, 0) has index 8 * leaf_lo into leaves."""
    samples, origins = [], []
    for a in range(nleaf[0]):
        for b in range(nleaf[1]):
            for c in range(nleaf[2]):
                blk = dense[8 * a:8 * a + 8, 8 * b:8 * b + 8, 8 * c:8 * c + 8]
                samples.append(np.ascontiguousarray(blk))  # C order: x * 64 + y * 8 + z
                origins.append([8 * (leaf_lo[0] + a), 8 * (leaf_lo[1] + b), 8 * (leaf_lo[2] + c)])
    return np.stack(samples).astype(np.float32), np.array(origins, np.int32)


def plane_dense(origin, h, leaf_lo, nleaf, d0=0.1):
    idx = np.stack(np.meshgrid(*[np.arange(8 * leaf_lo[k], 8 * (leaf_lo[k] + nleaf[k])) for k in range(3)],
                               indexing="ij"), axis=-1)
    return ((np.asarray(origin) + h * idx) @ N - d0).astype(np.float64)


def test_linear_function_reproduced():
    origin, h, lo, nl = (0.01, -0.02, 0.03), 0.01, (1, -2, 0), (3, 2, 2)
    dense = plane_dense(origin, h, lo, nl)
    sdf, org = leaves_of(dense.astype(np.float32), lo, nl)
    col = dict(shape=LEAVES, sdf=sdf, leaf_origins=org, origin=origin, spacing=h, params=(0.05,))
    rng = np.random.default_rng(3)
    lo_w = np.asarray(origin) + h * 8 * np.asarray(lo)
    ext = h * (8 * np.asarray(nl) - 1)
    for X in lo_w + ext * rng.random((300, 3)):
        phi, g = O.shape_phi(col, X)
        assert abs(phi - (X @ N - 0.1)) < 1e-6
        assert np.abs(g - N).max() < 1e-4


def test_tiled_box_equals_dense_grid():
    origin, h, lo, nl = (0.0, 0.0, 0.0), 1.0 / 64, (0, 0, 0), (2, 3, 2)
    rng = np.random.default_rng(4)
    dense = rng.standard_normal((16, 24, 16)).astype(np.float32)
    sdf, org = leaves_of(dense, lo, nl)
    leaves = dict(shape=LEAVES, sdf=sdf, leaf_origins=org, origin=origin, spacing=h, params=(1.0,))
    grid = dict(shape=GRID, sdf=dense, origin=origin, spacing=h)
    for X in (h * (np.array([15.0, 23.0, 15.0]) - 1e-3) * rng.random((400, 3))):
        p1, g1 = O.shape_phi(leaves, X)
        p2, g2 = O.shape_phi(grid, X)
        assert p1 == p2 and np.array_equal(g1, g2)
    # shuffled leaves (samples moved with their origins) give the same field
    perm = rng.permutation(len(org))
    shuf = dict(leaves, sdf=sdf[perm], leaf_origins=org[perm])
    for X in (h * 14.0 * rng.random((100, 3))):
        assert O.shape_phi(shuf, X)[0] == O.shape_phi(leaves, X)[0]


def test_negative_origins():
    origin, h, lo, nl = (0.5, 0.5, 0.5), 0.02, (-2, -1, -3), (2, 2, 3)
    dense = plane_dense(origin, h, lo, nl, d0=0.4)
    sdf, org = leaves_of(dense.astype(np.float32), lo, nl)
    col = dict(shape=LEAVES, sdf=sdf, leaf_origins=org, origin=origin, spacing=h, params=(0.05,))
    rng = np.random.default_rng(5)
    lo_w = np.asarray(origin) + h * 8 * np.asarray(lo)
    for X in lo_w + h * (8 * np.asarray(nl) - 1) * rng.random((200, 3)):
        phi, g = O.shape_phi(col, X)
        assert abs(phi - (X @ N - 0.4)) < 1e-6 and np.abs(g - N).max() < 1e-4


def test_background_outside_leaves():
    rng = np.random.default_rng(6)
    h, bg = 0.1, 0.75
    org = np.array([[0, 0, 0], [8, 0, 0], [0, 16, 8], [-8, 8, 0]], np.int32)
    sdf = rng.uniform(-1, 1, (4, 8, 8, 8)).astype(np.float32)
    col = dict(shape=LEAVES, sdf=sdf, leaf_origins=org, origin=(0.0, 0.0, 0.0), spacing=h, params=(bg,))
    phi, g = O.shape_phi(col, np.array([5.0, 5.0, 5.0]))  # cell (50, 50, 50): no leaf near
    assert phi == np.float32(bg) and not g.any()
    samples = {}
    for l, o in enumerate(org):
        for i in range(8):
            for j in range(8):
                for k in range(8):
                    samples[(o[0] + i, o[1] + j, o[2] + k)] = float(sdf[l, i, j, k])

    def brute(X):
        u = X / float(np.float32(h))  # the collider stores spacing as float32
        i0 = np.floor(u).astype(int)
        f = u - i0
        phi = 0.0
        for a in (0, 1):
            for b in (0, 1):
                for e in (0, 1):
                    w = (f[0] if a else 1 - f[0]) * (f[1] if b else 1 - f[1]) * (f[2] if e else 1 - f[2])
                    phi += w * samples.get((i0[0] + a, i0[1] + b, i0[2] + e), float(np.float32(bg)))
        return phi

    # cells straddling leaf faces and edges (some corners in no leaf), and cells inside leaves
    pts = [np.array(p) * h for p in ([7.5, 3.2, 7.7], [15.4, 7.6, 2.0], [-0.5, 8.3, 0.4], [3.3, 15.5, 8.6],
                                     [-8.7, 15.9, 7.2], [2.5, 2.5, 2.5], [7.99, 7.99, 7.99], [-3.0, 12.0, 4.0])]
    pts += list(h * rng.uniform(-10, 18, (300, 3)))
    for X in pts:
        assert abs(O.shape_phi(col, X)[0] - brute(X)) < 1e-12


def test_collision_uses_leaves():
    """A slip collider from leaves of a plane: inside nodes lose the normal velocity component."""
    origin, h, lo, nl = (0.0, 0.0, 0.0), 1.0 / 32, (0, 0, 0), (4, 4, 4)
    dense = plane_dense(origin, h, lo, nl, d0=0.3)
    sdf, org = leaves_of(dense.astype(np.float32), lo, nl)
    col = dict(type=1, shape=LEAVES, sdf=sdf, leaf_origins=org, origin=origin, spacing=h, params=(0.1,))
    x = np.array([0.4, 0.1, 0.3])
    assert x @ N - 0.3 < 0
    v, inside = O.resolve_collision(col, x, [1.0, 2.0, -1.0])
    vt = np.array([1.0, 2.0, -1.0]) - (np.array([1.0, 2.0, -1.0]) @ N) * N
    assert inside and np.allclose(v, vt, atol=1e-5)
