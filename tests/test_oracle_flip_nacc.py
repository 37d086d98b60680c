"""Oracle checks for the non-affine FLIP transfer, the Cam-clay return mapping and
ComputeParticleBeta (CPU only).

Each check is a property the mathematics fixes, computed by a route other than the
oracle's own code: the linear-reproduction identity of the B-spline gradients,
finite differences of the weights, the yield surface re-evaluated from the
Kirchhoff stress of the projected F (the separately pinned StVK-Hencky stress),
idempotence, rotation equivariance, and closed-form collider cases.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

DX = 1.0 / 64


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


# ---------------------------------------------------------------- FLIP weight gradients
def test_weight_gradient_reproduces_linear_fields():
    """sum_i x_i grad W_i^T = I and sum_i grad W_i = 0 (the quadratic B-spline
    reproduces linear functions), so the FLIP velocity gradient of an affine grid
    field b + A x_i is exactly A.  A sign or index error in dw fails this."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        xp = rng.uniform(0.3, 0.7, 3)
        base, W, G = O.weight_grad(xp, DX)
        X = np.zeros((3, 3, 3, 3))
        for ox in range(3):
            for oy in range(3):
                for oz in range(3):
                    X[ox, oy, oz] = (base + np.array([ox, oy, oz])) * DX
        assert abs(W.sum() - 1.0) < 1e-12
        np.testing.assert_allclose(G.reshape(27, 3).sum(0), 0.0, atol=1e-9)
        ident = np.einsum("abci,abcj->ij", X, G)
        np.testing.assert_allclose(ident, np.eye(3), atol=1e-9)
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        vel = X @ A.T + b
        np.testing.assert_allclose(np.einsum("abcr,abcj->rj", vel, G), A, atol=1e-8)


def test_weight_gradient_matches_finite_differences():
    """grad W_i = d W_i / d x_p by central differences of the weights (same base node)."""
    rng = np.random.default_rng(1)
    h = 1e-7
    for _ in range(20):
        xp = rng.uniform(0.3, 0.7, 3)
        # stay away from a base change within +-h
        f = xp / DX - np.floor(xp / DX - 0.5)
        if np.any(np.abs(f - 0.5) < 1e-3) or np.any(np.abs(f - 1.5) < 1e-3):
            continue
        base, W, G = O.weight_grad(xp, DX)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            b1, W1, _ = O.weight_grad(xp + e, DX)
            b2, W2, _ = O.weight_grad(xp - e, DX)
            assert (b1 == base).all() and (b2 == base).all()
            np.testing.assert_allclose((W1 - W2) / (2 * h), G[..., k], rtol=1e-5, atol=1e-3)


# ---------------------------------------------------------------- FLIP P2G and substep
def _one_particle(model, F, v, transfer):
    mats = [dict(model=model, E=2e4, nu=0.3, rho=1000.0, vol=DX ** 3 / 8, bulk=1e4)]
    x = np.array([[0.503], [0.497], [0.511]], np.float64)
    st = dict(x=x, v=np.asarray(v, np.float64).reshape(3, 1), C=np.zeros((9, 1)),
              F=np.asarray(F, np.float64).reshape(9, 1), mat=np.zeros(1, np.int32))
    s = O.Session(st, mats, DX, 1e-4, dtype=np.float64, boundary=O.BOUNDARY_NONE)
    s.set_transfer(transfer)
    return s, mats, x[:, 0]


def test_flip_p2g_single_particle():
    """FLIP P2G of one particle (Transfer.cu:892-933): grid m = W m, mv = W m v (no
    affine term), and the force channel W (-dt vol P F^T grad W) -- the reference's
    extra factor W -- node by node, with P from the separately pinned fcr_piola."""
    F = np.array([[1.02, 0.03, 0.0], [-0.01, 0.97, 0.02], [0.0, 0.01, 1.01]])
    v = [0.3, -0.2, 0.1]
    s, mats, xp = _one_particle(O.MODEL_FCR, F, v, O.TRANSFER_FLIP)
    s.stage("partition")
    s.stage("p2g")
    keys, data = s.grid()
    mu, lam = lame(2e4, float(np.float32(0.3)))  # the material table holds nu as float32
    P = O.fcr_piola(F, mu, lam)
    vol = mats[0]["vol"]
    mass = 1000.0 * vol
    contrib = -1e-4 * vol * P @ F.T
    base, W, G = O.weight_grad(xp, DX)
    nodes = {}
    for g, k in enumerate(keys):
        for c in range(64):
            lc = ((c >> 4) & 3, (c >> 2) & 3, c & 3)
            nodes[tuple(int(k[a]) + lc[a] for a in range(3))] = data[g, :, c]
    star = {}
    gs = s.gstar()
    for g, k in enumerate(keys):
        for c in range(64):
            lc = ((c >> 4) & 3, (c >> 2) & 3, c & 3)
            star[tuple(int(k[a]) + lc[a] for a in range(3))] = gs[g, :, c]
    for ox in range(3):
        for oy in range(3):
            for oz in range(3):
                key = tuple(int(base[a]) + o for a, o in enumerate((ox, oy, oz)))
                nd = nodes[key]
                w = W[ox, oy, oz]
                assert nd[0] == pytest.approx(w * mass, rel=1e-12)
                np.testing.assert_allclose(nd[1:], w * mass * np.asarray(v), rtol=1e-12, atol=1e-20)
                np.testing.assert_allclose(star[key][1:], w * (contrib @ G[ox, oy, oz]), rtol=1e-10, atol=1e-22)
    # no other node received anything
    touched = sum(1 for k, nd in nodes.items() if nd[0] != 0.0)
    assert touched == 27
    # the force channel summed over the nodes is contrib (1/2) grad sum W^2 (central differences)
    gsum = np.einsum("abc,abcj->j", W, G)
    h = 1e-7
    fd = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd[k] = ((O.weight_grad(xp + e, DX)[1] ** 2).sum() - (O.weight_grad(xp - e, DX)[1] ** 2).sum()) / (4 * h)
    np.testing.assert_allclose(gsum, fd, rtol=1e-5, atol=1e-6)
    tot = sum(star[k][1:] for k in star)
    np.testing.assert_allclose(tot, contrib @ fd, rtol=1e-4, atol=1e-18)


def _block(transfer, v0, E=0.0, n_cells=4, seed=3):
    rng = np.random.default_rng(seed)
    cells = np.stack(np.meshgrid(*[np.arange(28, 28 + n_cells)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = ((np.repeat(cells, 4, 0) + rng.uniform(0, 1, (len(cells) * 4, 3))) * DX).T
    n = x.shape[1]
    F = np.zeros((9, n))
    F[[0, 4, 8]] = 1.0
    st = dict(x=x, v=np.tile(np.asarray(v0, np.float64).reshape(3, 1), (1, n)), C=np.zeros((9, n)), F=F,
              mat=np.zeros(n, np.int32))
    mats = [dict(model=O.MODEL_FCR, E=E, nu=0.3, rho=1000.0, vol=DX ** 3 / 4)]
    s = O.Session(st, mats, DX, 1e-4, dtype=np.float64, boundary=O.BOUNDARY_NONE)
    s.set_transfer(transfer)
    return s, st


def test_flip_free_fall_of_stress_free_block():
    """Zero stiffness, uniform velocity: every FLIP substep gives v = v0 + k g dt,
    x += vstar dt, F = I and C untouched (zero)."""
    v0 = (0.4, -0.1, 0.2)
    s, st = _block(O.TRANSFER_FLIP, v0)
    k = 5
    s.substeps(k)
    out = s.state()
    vexp = np.asarray(v0) + k * 1e-4 * np.array([0.0, -9.8, 0.0])
    np.testing.assert_allclose(out["v"], np.tile(vexp[:, None], (1, out["v"].shape[1])), atol=1e-12)
    disp = sum(np.asarray(v0) + (j + 1) * 1e-4 * np.array([0.0, -9.8, 0.0]) for j in range(k)) * 1e-4
    np.testing.assert_allclose(out["x"] - st["x"], np.tile(disp[:, None], (1, out["x"].shape[1])), atol=1e-12)
    np.testing.assert_allclose(out["F"][[0, 4, 8]], 1.0, atol=1e-12)
    np.testing.assert_allclose(out["F"][[1, 2, 3, 5, 6, 7]], 0.0, atol=1e-12)
    assert np.all(out["C"] == 0.0)


def test_flip_elastic_block_stays_bounded_and_differs_from_aflip():
    """An elastic block under FLIP stays finite and near its start, and FLIP and AFLIP
    (which share the grid update) give different states (the transfers differ)."""
    s1, st = _block(O.TRANSFER_FLIP, (0.0, 0.0, 0.0), E=5e4)
    s2, _ = _block(O.TRANSFER_AFLIP, (0.0, 0.0, 0.0), E=5e4)
    s1.substeps(10)
    s2.substeps(10)
    a, b = s1.state(), s2.state()
    assert np.isfinite(a["x"]).all() and np.abs(a["x"] - st["x"]).max() < 0.1 * DX
    assert np.abs(a["F"] - b["F"]).max() > 0.0


# ---------------------------------------------------------------- Cam-clay
MU, LAM = lame(2e4, 0.3)
K = LAM + 2 * MU / 3


def _yield(F, logJp_in, phi=35.0, beta=2.0, xi=1.0):
    tau = O.stvk_hencky_tau(F, MU, LAM)
    p = -np.trace(tau) / 3
    dev = tau + p * np.eye(3)
    q2 = 1.5 * (dev * dev).sum()
    sp = math.sin(math.radians(phi))
    M = 6 * sp / (3 - sp)
    p0 = K * (1e-5 + math.sinh(xi * max(-logJp_in, 0.0)))
    return (1 + 2 * beta) * q2 + M * M * (p + beta * p0) * (p - p0), p, dev, p0


def _rand_rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_cam_clay_inside_unchanged():
    rng = np.random.default_rng(5)
    for _ in range(50):
        F = np.eye(3) + rng.normal(scale=1e-4, size=(3, 3))
        lj = -0.5
        y, *_ = _yield(F, lj)
        assert y < 0
        Fn, ljn, ch = O.project_cam_clay(F, MU, LAM, logJp=lj)
        assert not ch and ljn == lj
        np.testing.assert_array_equal(Fn, F)


def test_cam_clay_projects_onto_the_yield_surface_at_constant_pressure():
    """Middle branch: the projected stress has the trial pressure, the trial deviatoric
    direction, y = 0, and logJp unchanged."""
    rng = np.random.default_rng(6)
    hits = 0
    for _ in range(300):
        R = _rand_rot(rng)
        s = np.exp(rng.normal(scale=0.05, size=3))
        F = R @ np.diag(s) @ _rand_rot(rng).T
        lj = -rng.uniform(0.0, 0.2)
        y, p, dev, p0 = _yield(F, lj)
        if not (y > 0 and -2.0 * p0 < p < p0):
            continue
        hits += 1
        Fn, ljn, ch = O.project_cam_clay(F, MU, LAM, logJp=lj)
        assert ch
        yn, pn, devn, _ = _yield(Fn, lj)
        assert pn == pytest.approx(p, rel=1e-9, abs=1e-9 * K)
        assert abs(yn) < 1e-8 * K * K
        cosang = (dev * devn).sum() / (np.linalg.norm(dev) * np.linalg.norm(devn))
        assert cosang == pytest.approx(1.0, abs=1e-9)
        assert ljn == pytest.approx(lj, abs=1e-12)
        # idempotent
        F2, lj2, ch2 = O.project_cam_clay(Fn, MU, LAM, logJp=ljn)
        assert (not ch2) or np.abs(F2 - Fn).max() < 1e-10
    assert hits > 30


def test_cam_clay_tips_and_hardening():
    """Trial pressure beyond p0: the stress goes to -p0 I (compression tip) and logJp drops by
    the removed volume log J_trial - log J_new; beyond -beta p0: to beta p0 I."""
    lj = -0.01
    # compression
    F = np.diag([0.95, 0.96, 0.97])
    y, p, _, p0 = _yield(F, lj)
    assert y > 0 and p > p0
    Fn, ljn, ch = O.project_cam_clay(F, MU, LAM, logJp=lj)
    tau = O.stvk_hencky_tau(Fn, MU, LAM)
    np.testing.assert_allclose(tau, -p0 * np.eye(3), atol=1e-8 * K)
    assert ljn == pytest.approx(lj + math.log(np.linalg.det(F)) - math.log(np.linalg.det(Fn)), abs=1e-12)
    assert ljn < lj
    # tension
    F = np.diag([1.04, 1.05, 1.03])
    y, p, _, p0 = _yield(F, lj)
    assert y > 0 and p < -2.0 * p0
    Fn, ljn, ch = O.project_cam_clay(F, MU, LAM, logJp=lj)
    tau = O.stvk_hencky_tau(Fn, MU, LAM)
    np.testing.assert_allclose(tau, 2.0 * p0 * np.eye(3), atol=1e-8 * K)
    assert ljn > lj


def test_cam_clay_rotation_equivariance():
    rng = np.random.default_rng(7)
    for _ in range(20):
        F = np.eye(3) + rng.normal(scale=0.05, size=(3, 3))
        Q = _rand_rot(rng)
        a, la, _ = O.project_cam_clay(F, MU, LAM, logJp=-0.05)
        b, lb, _ = O.project_cam_clay(Q @ F, MU, LAM, logJp=-0.05)
        np.testing.assert_allclose(Q @ a, b, atol=1e-10)
        assert la == pytest.approx(lb, abs=1e-12)


# ---------------------------------------------------------------- ComputeParticleBeta
def test_particle_beta_cases():
    """Half-space y < 0.5 (phi = y - 0.5, normal +y): a particle inside moving down (into it) gets 0,
    inside moving up 0.5 when det F > 1, outside 0.5 / 0 by det F against jpc."""
    x = np.array([[0.5, 0.5, 0.5, 0.5], [0.49, 0.49, 0.6, 0.6], [0.5, 0.5, 0.5, 0.5]])
    v = np.array([[0, 0, 0, 0], [-1.0, 1.0, 0.0, 0.0], [0, 0, 0, 0]])
    F = np.zeros((9, 4))
    F[[0, 4, 8]] = 1.0
    F[0, 1] = 1.02   # det 1.02, inside moving up
    F[0, 2] = 1.02   # outside, det > 1
    F[0, 3] = 0.98   # outside, det < 1
    mats = [dict(model=O.MODEL_FCR, E=1e4, nu=0.3, rho=1000.0, vol=DX ** 3 / 8)]
    s = O.Session(dict(x=x, v=v, C=np.zeros((9, 4)), F=F, mat=np.zeros(4, np.int32)), mats, DX, 1e-4,
                  dtype=np.float64)
    s.set_colliders([dict(type=0, shape=0, params=(0.0, 1.0, 0.0, 0.5))])
    s.compute_particle_beta(1e-3, 1.0)
    beta, _ = s.aux()
    np.testing.assert_array_equal(beta, [0.0, 0.5, 0.5, 0.0])
    # a moving collider: the same particle at rest inside a half-space rising faster is "moving in"
    s.set_colliders([dict(type=0, shape=0, params=(0.0, 1.0, 0.0, 0.5), translation_rate=(0.0, 2.0, 0.0))])
    s.compute_particle_beta(1e-3, 1.0)
    beta, _ = s.aux()
    assert beta[1] == 0.0


def test_particle_beta_default_jpc_per_material():
    """jp_critical None: the reference's JpcDefault, 1 for every material, Cam-clay included (its
    1.1 branch, Boundary.cu:389-391, tests the anisotropic-model variant, which never holds
    Cam-clay): det F = 1.05 and 1.15 give 0.5 for both materials; an explicit jpc = 1.1 applies to
    every material and gives 0 at det F = 1.05."""
    x = np.full((3, 4), 0.5)
    F = np.zeros((9, 4))
    F[[0, 4, 8]] = 1.0
    F[0] = [1.05, 1.05, 1.15, 1.15]
    mat = np.array([0, 1, 0, 1], np.int32)
    mats = [dict(model=O.MODEL_STVK_HENCKY, E=1e4, nu=0.3, rho=1000.0, vol=DX ** 3 / 8),
            dict(model=O.MODEL_STVK_HENCKY, E=1e4, nu=0.3, rho=1000.0, vol=DX ** 3 / 8,
                 plasticity=O.PLASTIC_CAM_CLAY)]
    s = O.Session(dict(x=x, v=np.zeros((3, 4)), C=np.zeros((9, 4)), F=F, mat=mat), mats, DX, 1e-4,
                  dtype=np.float64)
    s.compute_particle_beta(1e-3)
    beta, _ = s.aux()
    np.testing.assert_array_equal(beta, [0.5, 0.5, 0.5, 0.5])
    s.compute_particle_beta(1e-3, 1.1)
    beta, _ = s.aux()
    np.testing.assert_array_equal(beta, [0.0, 0.0, 0.5, 0.5])
