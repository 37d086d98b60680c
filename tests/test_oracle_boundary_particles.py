"""Oracle checks for boundary particles (ZSBoundaryPrimitiveToZSGrid, Transfer.cu:1533-1622,
with ApplyGridBoundaryOnZSGrid, Boundary.cu:18-74), CPU only, by closed forms: a boundary
particle on a node gives that node its velocity with weight 1 and touches no other node; two
particles give the mass-weighted mean with the linear weights 1 - f, f; slip and separate
remove exactly the normal part of the relative velocity (separate only when approaching)."""
import numpy as np
import pytest

from oracle import oracle as O

DX = 1.0 / 32


def session(v0=(0.5, -0.25, 0.1)):
    rng = np.random.default_rng(2)
    cells = np.stack(np.meshgrid(*[np.arange(12, 18)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = ((np.repeat(cells, 2, 0) + rng.uniform(0, 1, (len(cells) * 2, 3))) * DX).T
    n = x.shape[1]
    F = np.zeros((9, n))
    F[[0, 4, 8]] = 1.0
    st = dict(x=x, v=np.tile(np.asarray(v0).reshape(3, 1), (1, n)), C=np.zeros((9, n)), F=F,
              mat=np.zeros(n, np.int32))
    mats = [dict(model=O.MODEL_FCR, E=0.0, nu=0.3, rho=1000.0, vol=DX ** 3 / 2)]
    return O.Session(st, mats, DX, 1e-4, dtype=np.float64, boundary=O.BOUNDARY_NONE)


def grid_after_boundary(s):
    for st in ("partition", "p2g", "grid_update", "boundary"):
        s.stage(st)
    keys, data = s.grid()
    out = {}
    for g, k in enumerate(keys):
        for c in range(64):
            if data[g, 0, c] != 0.0:
                out[(int(k[0]) + (c >> 4), int(k[1]) + ((c >> 2) & 3), int(k[2]) + (c & 3))] = data[g, 1:, c].copy()
    return out


def test_particle_on_a_node_sets_only_that_node():
    ref = grid_after_boundary(session())
    s = session()
    node = (15, 14, 16)
    vb = np.array([[1.5], [0.0], [-2.0]])
    s.set_boundary_particles(np.array(node, float).reshape(3, 1) * DX, vb, np.array([0.3]), type=0)
    got = grid_after_boundary(s)
    assert set(got) == set(ref)
    for k in ref:
        if k == node:
            np.testing.assert_allclose(got[k], vb[:, 0], rtol=1e-6)
        else:
            np.testing.assert_array_equal(got[k], ref[k])


def test_two_particles_give_the_mass_weighted_mean():
    s = session()
    node = np.array([14, 15, 15])
    # particle A a quarter cell past the node on x, particle B on the node
    xa = (node + np.array([0.25, 0.0, 0.0])) * DX
    xb = node * DX
    va, vb = np.array([1.0, 0.0, 0.0]), np.array([0.0, 2.0, 0.0])
    ma, mb = 0.2, 0.6
    s.set_boundary_particles(np.stack([xa, xb], 1), np.stack([va, vb], 1), np.array([ma, mb]), type=0)
    got = grid_after_boundary(s)
    wa = 0.75  # linear weight of A on the node below it: 1 - f, f = 1/4
    exp = (wa * ma * va + mb * vb) / (wa * ma + mb)
    np.testing.assert_allclose(got[tuple(node)], exp, rtol=1e-6)
    # the node past A gets only A: f = 1/4 -> weight 1/4, velocity va
    np.testing.assert_allclose(got[tuple(node + np.array([1, 0, 0]))], va, rtol=1e-6)


@pytest.mark.parametrize("typ", [1, 2])
def test_slip_and_separate_remove_the_normal_part(typ):
    v0 = np.array([0.5, -0.25, 0.1])
    ref = grid_after_boundary(session(tuple(v0)))
    node = (15, 15, 15)
    vobj = np.array([0.1, 0.3, 0.0])
    n = np.array([0.0, 0.6, 0.8])
    for sign in (1.0, -1.0):  # approaching (v - vobj) . n < 0, or leaving
        s = session(tuple(v0))
        nn = sign * n
        s.set_boundary_particles(np.array(node, float).reshape(3, 1) * DX, vobj.reshape(3, 1), np.array([1.0]),
                                 nrm=nn.reshape(3, 1), type=typ)
        got = grid_after_boundary(s)
        r = ref[node] - vobj
        proj = r @ nn
        exp = ref[node].copy() if (typ == 2 and proj >= 0) else r - proj * nn + vobj
        np.testing.assert_allclose(got[node], exp, rtol=1e-6, atol=1e-9)


def test_boundary_outside_the_grid_changes_nothing():
    ref = grid_after_boundary(session())
    s = session()
    s.set_boundary_particles(np.array([[0.9], [0.9], [0.9]]), np.ones((3, 1)), np.array([1.0]), type=0)
    got = grid_after_boundary(s)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k])


def test_particle_beta_boundary_grid_branch():
    """ComputeParticleBeta with a boundary grid (Boundary.cu:173-219): a particle whose linear
    stencil at x + dt v reaches boundary mass and moves into the boundary (v - v_b) . n_b < 0
    gets 0; the same particle moving out gets 0.5 / 0 by det F against jpc."""
    s = session()
    st = s.state()
    n = st["x"].shape[1]
    # a horizontal boundary sheet through the block at y = 14.5 cells, normal +y, at rest
    g = np.arange(11, 19, 0.5) * DX
    X, Z = np.meshgrid(g, g, indexing="ij")
    bx = np.stack([X.ravel(), np.full(X.size, 14.5 * DX), Z.ravel()])
    nb = bx.shape[1]
    nrm = np.tile(np.array([[0.0], [1.0], [0.0]]), (1, nb))
    s.set_boundary_particles(bx, np.zeros((3, nb)), np.full(nb, 0.1), nrm=nrm, type=1)
    s.compute_particle_beta(1e-4, 1.0)
    beta, _ = s.aux()
    y = st["x"][1] / DX
    near = (y > 13.6) & (y < 15.4)  # the linear stencil of x + dt v reaches the sheet's nodes 14, 15
    far = (y < 12.9) | (y > 16.1)
    # the block moves down (v_y = -0.25) into the sheet: moving in near it, det F = 1 (not > 1) elsewhere
    assert np.all(beta[near] == 0.0)
    assert np.all(beta[far] == 0.0)
    # moving up: nobody moves in; det F = 1 is not above jpc = 1, but is above jpc = 0.99
    s2 = session(v0=(0.5, 0.25, 0.1))
    s2.set_boundary_particles(bx, np.zeros((3, nb)), np.full(nb, 0.1), nrm=nrm, type=1)
    s2.compute_particle_beta(1e-4, 0.99)
    beta2, _ = s2.aux()
    assert np.all(beta2 == 0.5)
    s.compute_particle_beta(1e-4, 0.99)
    beta3, _ = s.aux()
    assert np.all(beta3[near] == 0.0) and np.all(beta3[far] == 0.5)
