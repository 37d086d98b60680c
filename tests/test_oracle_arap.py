"""Oracle checks for the anisotropic ARAP fiber term ("aniso" = "arap", AnisotropicArap{E, nu,
strength}, Generation.cu:60-63, added to P at Transfer.cu:866-867; zpc's AnisotropicArap is
absent, so the published energy is the definition): Kim, De Goes, Iben 2019,
  psi = mu_a/2 (sqrt(I5) - S(I4))^2, I5 = |F a|^2, I4 = a^T S a (F = R S, R proper), S(x) = sign.
Pinned by FP64 central differences of psi (the polar taken here from numpy's SVD), the
inversion-safe closed form along the fiber P = mu_a (lam - 1) a a^T for F = I + (lam - 1) a a^T
at any lam != 0 (both signs of I4), rest stability under rotations and stretches across the
fiber, and, through P2G, the APIC inertia identity sum_i dp_i d_i^T = -dt vol P_a F^T."""
import numpy as np
import pytest

from oracle import oracle as O

MUA = 2.5e4
DX = 1.0 / 32


def polar_R(F):
    U, s, Vt = np.linalg.svd(F)
    R = U @ Vt
    if np.linalg.det(R) < 0:  # signed SVD: flip the smallest singular value
        U = U.copy()
        U[:, 2] = -U[:, 2]
        R = U @ Vt
    return R


def psi(F, a):
    R = polar_R(F)
    I4 = a @ (R.T @ F) @ a
    I5 = (F @ a) @ (F @ a)
    return MUA / 2 * (np.sqrt(I5) - (-1.0 if I4 < 0 else 1.0)) ** 2


def rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@pytest.mark.parametrize("inverted", [False, True])
def test_fiber_piola_is_the_energy_gradient(inverted):
    rng = np.random.default_rng(5 + inverted)
    h = 1e-7
    done = 0
    while done < 25:
        F = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        if inverted:
            F[:, 0] = -F[:, 0]
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        J = np.linalg.det(F)
        R = polar_R(F)
        I4 = a @ (R.T @ F) @ a
        sv = np.linalg.svd(F, compute_uv=False)
        # stay away from the sign switch of I4 and from degenerate singular values
        if abs(J) < 0.1 or abs(I4) < 0.05 or np.diff(sv).__abs__().min() < 0.05 or (J > 0) == inverted:
            continue
        P = O.arap_fiber_piola(F, a, MUA)
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                G[i, j] = (psi(F + E, a) - psi(F - E, a)) / (2 * h)
        np.testing.assert_allclose(P, G, rtol=1e-5, atol=1e-5 * MUA)
        done += 1


@pytest.mark.parametrize("lam", [1.3, 0.7, 0.2, -0.4, -0.9])
def test_uniaxial_along_the_fiber_is_inversion_safe(lam):
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        F = np.eye(3) + (lam - 1.0) * np.outer(a, a)
        Q = rot(rng)
        # I + (lam - 1) a a^T and Q times it: the fiber term rotates with F
        np.testing.assert_allclose(O.arap_fiber_piola(F, a, MUA), MUA * (lam - 1.0) * np.outer(a, a),
                                   atol=1e-9 * MUA)
        np.testing.assert_allclose(O.arap_fiber_piola(Q @ F, a, MUA), Q @ (MUA * (lam - 1.0) * np.outer(a, a)),
                                   atol=1e-9 * MUA)


def test_rest_and_stretch_across_the_fiber():
    rng = np.random.default_rng(8)
    for _ in range(10):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        Q = rot(rng)
        np.testing.assert_allclose(O.arap_fiber_piola(Q, a, MUA), 0.0, atol=1e-9 * MUA)
        # stretch across the fiber: F a = Q a keeps unit length
        b = np.cross(a, rng.normal(size=3))
        b /= np.linalg.norm(b)
        F = Q @ (np.eye(3) + 0.4 * np.outer(b, b))
        np.testing.assert_allclose(O.arap_fiber_piola(F, a, MUA), 0.0, atol=1e-9 * MUA)
    # zero stiffness, zero term
    np.testing.assert_array_equal(O.arap_fiber_piola(np.diag([1.2, 0.9, 1.1]), np.array([0.0, 1.0, 0.0]), 0.0), 0.0)


def one_particle(aniso, F, a, strength=10.0):
    xp = np.array([[15.37], [16.81], [14.42]]) * DX
    st = dict(x=xp, v=np.zeros((3, 1)), C=np.zeros((9, 1)), F=F.reshape(9, 1), mat=np.zeros(1, np.int32))
    mats = [dict(model=O.MODEL_FCR, E=1e4, nu=0.3, rho=1000.0, vol=DX ** 3 / 4, aniso=aniso,
                 aniso_strength=strength)]
    s = O.Session(st, mats, DX, 1e-4, dtype=np.float64, boundary=O.BOUNDARY_NONE)
    s.set_fiber(a.reshape(3, 1))
    s.stage("partition")
    s.stage("p2g")
    keys, data = s.grid()
    nodes, mom = [], []
    for g, k in enumerate(keys):
        for c in range(64):
            if data[g, 0, c] != 0.0:
                nodes.append((int(k[0]) + (c >> 4), int(k[1]) + ((c >> 2) & 3), int(k[2]) + (c & 3)))
                mom.append(data[g, 1:, c])
    order = np.lexsort(np.array(nodes).T)
    return np.array(nodes)[order], np.array(mom)[order], xp[:, 0], mats[0]


def test_p2g_adds_the_fiber_stress():
    """One particle, P2G alone: the ARAP run minus the isotropic run, summed as
    sum_i dp_i (x_i - x_p)^T, is -dt vol P_a F^T (the quadratic B-spline's inertia D = dx^2/4 I
    times D^-1 = 4/dx^2 of the stress term), with P_a from the uniaxial closed form."""
    a = np.array([0.6, 0.0, 0.8])
    lam = 1.25
    F = np.eye(3) + (lam - 1.0) * np.outer(a, a)
    n0, m0, xp, mt = one_particle(O.ANISO_NONE, F, a)
    n1, m1, _, _ = one_particle(O.ANISO_ARAP, F, a, strength=3.0)
    assert np.array_equal(n0, n1) and len(n0) == 27
    d = n0 * DX - xp[None, :]
    M = ((m1 - m0)[:, :, None] * d[:, None, :]).sum(0)
    mu = mt["E"] / (2 * (1 + mt["nu"]))
    Pa = 3.0 * mu * (lam - 1.0) * np.outer(a, a)
    np.testing.assert_allclose(M, -1e-4 * mt["vol"] * Pa @ F.T, rtol=1e-6, atol=1e-12)
    # strength 0: the same grid as without the fiber model
    _, m2, _, _ = one_particle(O.ANISO_ARAP, F, a, strength=0.0)
    np.testing.assert_array_equal(m2, m0)
