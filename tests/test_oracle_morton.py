"""The oracle's Morton ordering (ZenoParticles::orderByMortonCode, projects/CUDA/Structures.hpp:406-443).

zpc's morton_code<3> is absent from the reference tree, so the restated code is
pinned by what the construction fixes rather than by zpc's own output:
- known answers: one set bit per axis lands at 3b + 2 (x), 3b + 1 (y), 3b (z);
- an independent formulation, the multiply-and-mask bit spreading of Karras
  2012 (which the GPU kernel uses), agrees with the bit-by-bit loop on random
  and extreme inputs;
- the code is strictly increasing in each quantised coordinate with the others
  fixed, and the per-axis quantisation truncates c * 1024 and clamps to
  [0, 1023];
- the order is a stable sort (equal codes keep the index order) and the two
  maps are inverse permutations.
"""
import numpy as np

from oracle import oracle as O


def spread10(v):
    """Karras 2012's bit spreading of a 10-bit integer to every third bit (numpy, uint64 to stay exact)."""
    v = v.astype(np.uint64)
    v = (v * np.uint64(0x00010001)) & np.uint64(0xFF0000FF)
    v = (v * np.uint64(0x00000101)) & np.uint64(0x0F00F00F)
    v = (v * np.uint64(0x00000011)) & np.uint64(0xC30C30C3)
    v = (v * np.uint64(0x00000005)) & np.uint64(0x49249249)
    return v


def magic_code(q):
    return ((spread10(q[0]) << np.uint64(2)) | (spread10(q[1]) << np.uint64(1)) | spread10(q[2])).astype(np.uint32)


def coords_of(q):
    """Coordinates whose quantisation is exactly q (cell centres of the 1024 grid)."""
    return ((q.astype(np.float64) + 0.5) / 1024.0).astype(np.float32)


def test_known_answers():
    q = np.array([[0, 1, 0, 0, 2, 1023, 5], [0, 0, 1, 0, 0, 1023, 3], [0, 0, 0, 1, 0, 1023, 6]], np.uint32)
    codes = O.morton_code3(coords_of(q))
    assert codes.tolist() == [0, 4, 2, 1, 32, (1 << 30) - 1,
                              # x = 101b, y = 011b, z = 110b -> bits (x2 y2 z2)(x1 y1 z1)(x0 y0 z0) = 101 011 110
                              0b101011110]


def test_matches_magic_number_spreading():
    rng = np.random.default_rng(7)
    q = rng.integers(0, 1024, size=(3, 20000)).astype(np.uint32)
    q[:, :8] = np.array([[0, 1023, 0, 0, 1023, 1023, 0, 512], [0, 1023, 1023, 0, 0, 1023, 1023, 512],
                         [0, 1023, 0, 1023, 0, 0, 1023, 512]], np.uint32)
    assert np.array_equal(O.morton_code3(coords_of(q)), magic_code(q))


def test_monotone_per_axis_and_quantisation():
    rng = np.random.default_rng(8)
    base = rng.integers(0, 1024, size=(3, 200)).astype(np.uint32)
    for axis in range(3):
        lo, hi = base.copy(), base.copy()
        lo[axis] = np.minimum(base[axis], 1022)
        hi[axis] = lo[axis] + 1
        assert np.all(O.morton_code3(coords_of(hi)) > O.morton_code3(coords_of(lo)))
    # truncation of c * 1024 and the clamp to [0, 1023]
    c = np.array([[-0.5, 0.0, 0.00097, 0.000977, 0.9999, 1.0, 3.0]] * 3, np.float32)
    q = np.minimum(np.maximum(np.floor(c.astype(np.float64) * 1024.0), 0.0), 1023.0).astype(np.uint32)
    assert np.array_equal(O.morton_code3(c), magic_code(q))


def test_uniform_coord_and_stable_order():
    rng = np.random.default_rng(9)
    n = 5000
    # many particles share a quantised cell (equal codes): the sort must keep their index order
    x = (0.2 + 0.01 * rng.integers(0, 8, size=(3, n)) + 1e-5 * rng.random((3, n))).astype(np.float32)
    lo, hi = (0.2, 0.2, 0.2), (0.3, 0.3, 0.3)
    c = O.uniform_coord(x, lo, hi)
    assert np.array_equal(c, ((x - np.float32(0.2)) / (np.float32(0.3) - np.float32(0.2))).astype(np.float32))
    codes, o2o, inv = O.order_by_morton_code(x, lo, hi)
    ordered = codes[o2o]
    assert np.all(np.diff(ordered.astype(np.int64)) >= 0)
    same = np.diff(ordered.astype(np.int64)) == 0
    assert same.sum() > n // 2  # the test does exercise ties
    assert np.all(np.diff(o2o)[same] > 0)
    assert np.array_equal(inv[o2o], np.arange(n))
    assert np.array_equal(o2o[inv], np.arange(n))
