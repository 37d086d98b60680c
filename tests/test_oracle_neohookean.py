"""Oracle checks for the Neo-Hookean ("nhk") and stable Neo-Hookean ("snhk") stresses
(ConfigConstitutiveModel, Generation.cu:43-48; zpc's models are absent, so the published
energies are the definition): P = dpsi/dF by FP64 central differences of
  psi_nhk  = mu/2 (|F|^2 - 3) - mu log J + lam/2 (log J)^2          (Jiang et al. 2016, eq. 48)
  psi_snhk = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2        (Smith, de Goes, Kim 2018)
rest stability P(I) = 0, isotropy P(QF) = Q P(F), and the small-strain limits."""
import numpy as np
import pytest

from oracle import oracle as O

MU, LAM = 1.3e4, 2.1e4


def psi(F, stable):
    J = np.linalg.det(F)
    I = (F * F).sum()
    if stable:
        return MU / 2 * (I - 3) - MU * (J - 1) + LAM / 2 * (J - 1) ** 2
    return MU / 2 * (I - 3) - MU * np.log(J) + LAM / 2 * np.log(J) ** 2


def rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@pytest.mark.parametrize("stable", [False, True])
def test_piola_is_the_energy_gradient(stable):
    rng = np.random.default_rng(11 + stable)
    h = 1e-6
    for _ in range(30):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        if np.linalg.det(F) < 0.2:
            continue
        P = O.neo_hookean_piola(F, MU, LAM, stable)
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                G[i, j] = (psi(F + E, stable) - psi(F - E, stable)) / (2 * h)
        np.testing.assert_allclose(P, G, rtol=1e-6, atol=1e-6 * MU)


@pytest.mark.parametrize("stable", [False, True])
def test_rest_stability_and_isotropy(stable):
    np.testing.assert_allclose(O.neo_hookean_piola(np.eye(3), MU, LAM, stable), 0.0, atol=1e-9)
    rng = np.random.default_rng(21)
    for _ in range(10):
        F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
        Q = rot(rng)
        np.testing.assert_allclose(O.neo_hookean_piola(Q @ F, MU, LAM, stable),
                                   Q @ O.neo_hookean_piola(F, MU, LAM, stable), atol=1e-8 * MU)


def test_small_strain_limits():
    """Neo-Hookean linearises to 2 mu eps + lam tr(eps) I; the stable form (Lame constants taken
    as given) to 2 mu eps + (lam - mu) tr(eps) I."""
    rng = np.random.default_rng(31)
    H = 1e-6 * rng.normal(size=(3, 3))
    eps = 0.5 * (H + H.T)
    lin = 2 * MU * eps + LAM * np.trace(eps) * np.eye(3)
    np.testing.assert_allclose(O.neo_hookean_piola(np.eye(3) + H, MU, LAM, False), lin, rtol=1e-4, atol=1e-3 * MU * 1e-6)
    lin_s = 2 * MU * eps + (LAM - MU) * np.trace(eps) * np.eye(3)
    np.testing.assert_allclose(O.neo_hookean_piola(np.eye(3) + H, MU, LAM, True), lin_s, rtol=1e-4,
                               atol=1e-3 * MU * 1e-6)


def test_inverted_f():
    """The stable form is defined for inverted F and pushes it back (P . F^-T direction
    restores J); Neo-Hookean clamps J at 1e-6 and stays finite."""
    F = np.diag([1.0, 1.0, -0.5])
    P = O.neo_hookean_piola(F, MU, LAM, True)
    assert np.isfinite(P).all()
    P2 = O.neo_hookean_piola(F, MU, LAM, False)
    assert np.isfinite(P2).all()
