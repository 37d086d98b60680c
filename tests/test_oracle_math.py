"""Pins the oracle's math primitives against closed forms, invariants and library routines.

The reference snapshot holds no golden vectors for this path (its zpc arithmetic
library is absent), so each primitive is pinned by what the mathematics fixes:
  - signed SVD / polar rotation  vs numpy.linalg.svd and scipy.linalg.polar (FP64);
  - fixed-corotated P            vs closed forms (P(R) = 0, diagonal F), isotropy and
                                  FP64 central differences of the energy;
  - B-spline arena               Sum w = 1, Sum w d = 0, Sum w d^2 = dx^2/4 per axis (these three
                                  moment conditions determine the three quadratic weights uniquely);
  - cell coordinates             FP32 product then floor, including exact half-cell boundaries;
  - StVK-Hencky tau              vs the diagonal closed form, isotropy, and P = tau F^-T against FP64
                                  central differences of psi = mu |log s|^2 + lam/2 (sum log s)^2 with s
                                  from numpy.linalg.svd;
  - return mappings              Drucker-Prager: tension resets F to its polar rotation (scipy),
                                  yielding states land on the cone with tr(e) kept and e^ direction
                                  kept, elastic states and projected states do not move, isotropy;
                                  von Mises: |e^'| = yield / (2 mu), tr(e) kept, the same fixed points.
"""
import numpy as np
import pytest
import scipy.linalg


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def test_svd3_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    for trial in range(400):
        F = np.eye(3) + 0.6 * rng.normal(size=(3, 3))
        if trial % 4 == 0:
            F[:, 0] *= -1  # inverted
        if trial % 7 == 0:
            F = rand_rot(rng) @ np.diag([1.3, 0.7, 1e-4]) @ rand_rot(rng)  # nearly degenerate
        U, s, V = oracle.svd3(F)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-12)
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-12)
        assert np.isclose(np.linalg.det(U), 1.0, atol=1e-12) and np.isclose(np.linalg.det(V), 1.0, atol=1e-12)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-11 * max(1, np.abs(F).max()))
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.abs(s), ref, rtol=1e-9, atol=1e-12)
        assert s[0] >= s[1] >= abs(s[2]) - 1e-14
        assert np.sign(s[2]) == np.sign(np.linalg.det(F)) or abs(s[2]) < 1e-14


def test_polar_rotation_matches_scipy(oracle):
    rng = np.random.default_rng(1)
    for trial in range(300):
        F = rand_rot(rng) @ np.diag(rng.uniform(0.3, 2.0, 3)) @ rand_rot(rng)
        R = oracle.polar_rotation(F)
        Rs, S = scipy.linalg.polar(F)  # F = Rs S, unique for det F > 0
        assert np.allclose(R, Rs, atol=1e-10)
        S2 = R.T @ F
        assert np.allclose(S2, S2.T, atol=1e-10)
        assert np.all(np.linalg.eigvalsh((S2 + S2.T) / 2) > 0)


def test_polar_rotation_inverted_convention(oracle):
    """det F < 0: R is the proper rotation U diag(1, 1, -1) V^T-corrected factor that flips the
    smallest singular value (Irving et al. 2004)."""
    rng = np.random.default_rng(2)
    for trial in range(200):
        F = rand_rot(rng) @ np.diag(rng.uniform(0.3, 2.0, 3) * np.array([1, 1, -1])) @ rand_rot(rng)
        R = oracle.polar_rotation(F)
        U0, s0, Vt0 = np.linalg.svd(F)
        D = np.diag([1.0, 1.0, np.linalg.det(U0 @ Vt0)])
        Rexp = U0 @ D @ Vt0
        assert np.isclose(np.linalg.det(R), 1.0, atol=1e-12)
        assert np.allclose(R, Rexp, atol=1e-9)


def fcr_energy(F, mu, lam):
    U0, s0, Vt0 = np.linalg.svd(F)
    R = U0 @ np.diag([1, 1, np.linalg.det(U0 @ Vt0)]) @ Vt0
    J = np.linalg.det(F)
    return mu * np.sum((F - R) ** 2) + 0.5 * lam * (J - 1) ** 2


def test_fcr_closed_forms(oracle):
    rng = np.random.default_rng(3)
    mu, lam = 2.1e4, 1.4e4
    for _ in range(50):
        R = rand_rot(rng)
        assert np.allclose(oracle.fcr_piola(R, mu, lam), 0.0, atol=1e-9)
        sig = rng.uniform(0.5, 1.6, 3)
        P = oracle.fcr_piola(np.diag(sig), mu, lam)
        J = sig.prod()
        Pexp = np.diag(2 * mu * (sig - 1) + lam * (J - 1) * J / sig)
        assert np.allclose(P, Pexp, rtol=1e-10, atol=1e-8)


def test_fcr_isotropy_and_energy_gradient(oracle):
    rng = np.random.default_rng(4)
    mu, lam = 3.0e4, 2.0e4
    for _ in range(40):
        F = rand_rot(rng) @ np.diag(rng.uniform(0.6, 1.5, 3)) @ rand_rot(rng)
        Q = rand_rot(rng)
        P = oracle.fcr_piola(F, mu, lam)
        assert np.allclose(oracle.fcr_piola(Q @ F, mu, lam), Q @ P, rtol=1e-9, atol=1e-7)
        assert np.allclose(oracle.fcr_piola(F @ Q, mu, lam), P @ Q, rtol=1e-9, atol=1e-7)
        h = 1e-6
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                G[i, j] = (fcr_energy(F + E, mu, lam) - fcr_energy(F - E, mu, lam)) / (2 * h)
        assert np.allclose(P, G, rtol=1e-5, atol=1e-3)


def test_fcr_f32_close_to_f64(oracle):
    rng = np.random.default_rng(5)
    mu, lam = oracle.lame_f32(5e4, 0.2)
    for _ in range(50):
        F = (rand_rot(rng) @ np.diag(rng.uniform(0.8, 1.2, 3)) @ rand_rot(rng)).astype(np.float32)
        P32 = oracle.fcr_piola(F, mu, lam, dtype=np.float32)
        P64 = oracle.fcr_piola(F.astype(np.float64), mu, lam)
        assert np.allclose(P32, P64, rtol=1e-4, atol=1e-2 * (mu + lam) * 1e-3)


def test_lame():
    from oracle import oracle as O
    for E, nu in [(5e4, 0.2), (1e5, 0.4), (1e4, 0.3)]:
        mu, lam = O.lame_f32(E, nu)
        assert np.isclose(mu, E / (2 * (1 + nu)), rtol=1e-6)
        assert np.isclose(lam, E * nu / ((1 + nu) * (1 - 2 * nu)), rtol=1e-6)


@pytest.mark.parametrize("dx", [1 / 64, 1 / 128, 1 / 256])
def test_arena_moments(oracle, dx):
    rng = np.random.default_rng(6)
    for _ in range(300):
        xp = rng.uniform(0.05, 0.95, 3)
        base, w, d = oracle.arena(xp, dx)
        X = xp / dx
        assert np.array_equal(base, np.floor(X - 0.5).astype(np.int32))
        for k in range(3):
            f = X[k] - base[k]
            assert 0.5 <= f < 1.5
            assert np.allclose(d[k], (np.arange(3) - f) * dx, atol=1e-15)
            assert np.isclose(w[k].sum(), 1.0, atol=1e-14)
            assert np.isclose((w[k] * d[k]).sum(), 0.0, atol=1e-14 * dx)
            assert np.isclose((w[k] * d[k] ** 2).sum(), dx * dx / 4, rtol=1e-12)
            assert np.all(w[k] >= 0)


def test_cell_coords_fp32_boundaries(oracle):
    dx = np.float32(1 / 128)
    dxinv = np.float32(1.0) / dx
    rng = np.random.default_rng(7)
    x = rng.uniform(-0.02, 1.02, size=(3, 5000)).astype(np.float32)
    # exact half-cell boundaries: X = k + 0.5 gives base k; one ulp below gives k - 1
    k = np.arange(-2, 130, dtype=np.float32)
    xb = ((k + np.float32(0.5)) / dxinv).astype(np.float32)
    x = np.concatenate([x, np.stack([xb, xb, xb]), np.stack([np.nextafter(xb, -1)] * 3)], axis=1)
    c = oracle.cell_coords(x, float(dx))
    X = (x * dxinv).astype(np.float32)  # FP32 product, as the reference's x * dxinv
    expected = np.floor(X.astype(np.float64) - 0.5).astype(np.int32)
    assert np.array_equal(c, expected)


# ------------------------------------------------------------------ StVK-Hencky and return mapping
def hencky_energy(F, mu, lam):
    s = np.linalg.svd(F, compute_uv=False)
    e = np.log(np.abs(s))
    return mu * np.sum(e * e) + 0.5 * lam * np.sum(e) ** 2


def hencky_strain(F):
    s = np.linalg.svd(F, compute_uv=False)
    return np.log(np.abs(s))


def test_stvk_hencky_closed_form_and_isotropy(oracle):
    mu, lam = 3.0e4, 5.0e4
    d = np.array([1.2, 0.9, 0.7])
    tau = oracle.stvk_hencky_tau(np.diag(d), mu, lam)
    e = np.log(d)
    assert np.allclose(tau, np.diag(2 * mu * e + lam * e.sum()), rtol=1e-12, atol=1e-9)
    rng = np.random.default_rng(3)
    for _ in range(50):
        F = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.05:
            continue
        Q, R = rand_rot(rng), rand_rot(rng)
        t0 = oracle.stvk_hencky_tau(F, mu, lam)
        assert np.allclose(oracle.stvk_hencky_tau(Q @ F @ R, mu, lam), Q @ t0 @ Q.T, rtol=1e-9, atol=1e-6)


def test_stvk_hencky_energy_gradient(oracle):
    """P = tau F^-T is dpsi/dF: FP64 central differences of the Hencky energy."""
    mu, lam = 2.0e4, 3.0e4
    rng = np.random.default_rng(4)
    h = 1e-6
    for _ in range(30):
        F = np.eye(3) + 0.25 * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.1:
            continue
        P = oracle.stvk_hencky_tau(F, mu, lam) @ np.linalg.inv(F).T
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                G[i, j] = (hencky_energy(F + E, mu, lam) - hencky_energy(F - E, mu, lam)) / (2 * h)
        assert np.allclose(P, G, rtol=1e-5, atol=1e-3 * np.abs(G).max())


def _dp_state(rng, tr, dev):
    """F = Q diag(exp(e)) R with tr(e) = tr and a deviatoric part of norm dev."""
    a = rng.normal(size=3)
    a -= a.mean()
    a *= dev / np.linalg.norm(a)
    e = a + tr / 3.0
    return rand_rot(rng) @ np.diag(np.exp(e)) @ rand_rot(rng), e


def test_drucker_prager_return_mapping(oracle):
    mu, lam, phi = 1.0e4, 1.5e4, 30.0
    sp = np.sin(np.radians(phi))
    alpha = np.sqrt(2.0 / 3.0) * 2 * sp / (3 - sp)
    c = (3 * lam + 2 * mu) / (2 * mu)
    rng = np.random.default_rng(5)
    # tension: the strain resets, F becomes its rotation factor
    F, _ = _dp_state(rng, 0.05, 0.02)
    Fp, changed = oracle.project_strain(F, mu, lam, oracle.PLASTIC_DRUCKER_PRAGER, phi)
    assert changed and np.allclose(Fp, scipy.linalg.polar(F)[0], atol=1e-12)
    for _ in range(40):
        tr = -rng.uniform(0.01, 0.2)
        cone = -c * tr * alpha  # |e^| on the yield surface
        # inside the cone: unchanged
        F, _ = _dp_state(rng, tr, 0.5 * cone)
        Fp, changed = oracle.project_strain(F, mu, lam, oracle.PLASTIC_DRUCKER_PRAGER, phi)
        assert not changed and np.array_equal(Fp, F)
        # outside: projected onto the cone, tr(e) and the direction of e^ kept
        F, e = _dp_state(rng, tr, 3.0 * cone + 0.01)
        Fp, changed = oracle.project_strain(F, mu, lam, oracle.PLASTIC_DRUCKER_PRAGER, phi)
        assert changed
        e1 = hencky_strain(Fp)
        assert abs(e1.sum() - tr) < 1e-10
        eh = np.sort(e1 - e1.mean())
        eh0 = np.sort(e - e.mean())
        assert np.isclose(np.linalg.norm(eh), cone, rtol=1e-9)
        assert np.allclose(eh / np.linalg.norm(eh), eh0 / np.linalg.norm(eh0), atol=1e-9)
        # the right factor is kept: F'^T F' and F^T F share eigenvectors, F' F^-1 is symmetric
        M = Fp @ np.linalg.inv(F)
        assert np.allclose(M, M.T, atol=1e-10)
        # a projected state is a fixed point; isotropy
        Fq, _ = oracle.project_strain(Fp, mu, lam, oracle.PLASTIC_DRUCKER_PRAGER, phi)
        assert np.allclose(Fq, Fp, atol=1e-12)
        Q, R = rand_rot(rng), rand_rot(rng)
        Fr, _ = oracle.project_strain(Q @ F @ R, mu, lam, oracle.PLASTIC_DRUCKER_PRAGER, phi)
        assert np.allclose(Fr, Q @ Fp @ R, atol=1e-10)


def test_von_mises_return_mapping(oracle):
    mu, lam, ys = 1.0e4, 1.5e4, 300.0
    r = ys / (2 * mu)
    rng = np.random.default_rng(6)
    for _ in range(40):
        tr = rng.uniform(-0.1, 0.1)
        F, _ = _dp_state(rng, tr, 0.8 * r)
        Fp, changed = oracle.project_strain(F, mu, lam, oracle.PLASTIC_VON_MISES, yield_stress=ys)
        assert not changed and np.array_equal(Fp, F)
        F, e = _dp_state(rng, tr, r * rng.uniform(1.5, 4.0))
        Fp, changed = oracle.project_strain(F, mu, lam, oracle.PLASTIC_VON_MISES, yield_stress=ys)
        assert changed
        e1 = hencky_strain(Fp)
        assert abs(e1.sum() - tr) < 1e-10
        assert np.isclose(np.linalg.norm(e1 - e1.mean()), r, rtol=1e-9)
        Fq, _ = oracle.project_strain(Fp, mu, lam, oracle.PLASTIC_VON_MISES, yield_stress=ys)
        assert np.allclose(Fq, Fp, atol=1e-12)
