"""Pins the oracle's J-based fluid equation of state (the "J" channel, Generation.cpp:878-884)
by what the mechanics fixes, not by its own formula (FP64 sessions).

  energy     psi(J) = K/2 (J - 1)^2 per reference volume is the weakly compressible fluid
             (Cauchy pressure p = -dpsi/dJ = -K (J - 1)); its Kirchhoff stress is
             tau = J dpsi/dJ I.  Through one particle's P2G alone, the B-spline inertia
             identity sum_i W d d^T = dx^2/4 I with D^-1 = 4/dx^2 turns the scattered momentum
             into sum_i dp_i (x_i - x_p)^T = -dt vol tau.  tau comes here from central
             differences of psi, so tau = K (J - 1) I (no J factor), a flipped sign or a
             missing volume fails.
  sign       a uniformly compressed block (J < 1) pushes its faces outward, an expanded one
             (J > 1) pulls them in, J = 1 moves nothing; total momentum stays zero.
  hydrostatics  a column started at J(depth d) = 1 - rho g d / K stays at rest under gravity
             (its velocities stay a small fraction of free fall), while the same column
             started at J = 1 falls and compresses.
"""
import numpy as np
import pytest

from oracle import oracle as O

DX = 1.0 / 32
K_BULK = 4.0e3


def fluid_state(x, J, v=None):
    n = x.shape[1]
    F = np.zeros((9, n))
    F[0] = J
    return dict(x=x, v=np.zeros((3, n)) if v is None else v, C=np.zeros((9, n)), F=F,
                mat=np.zeros(n, np.int32))


def one_particle_grid(J, dt=1e-4, vol=DX ** 3 / 8):
    xp = np.array([[15.37], [16.81], [14.42]]) * DX
    mats = [dict(model=O.MODEL_JFLUID, rho=1000.0, vol=vol, bulk=K_BULK)]
    s = O.Session(fluid_state(xp, J), mats, DX, dt, gravity=(0, 0, 0), boundary=O.BOUNDARY_NONE,
                  dtype=np.float64)
    s.stage("partition")
    s.stage("p2g")
    keys, data = s.grid()
    nodes, mom = [], []
    for g, k in enumerate(keys):
        for c in range(64):
            if data[g, 0, c] != 0.0:
                nodes.append((int(k[0]) + (c >> 4), int(k[1]) + ((c >> 2) & 3), int(k[2]) + (c & 3)))
                mom.append(data[g, 1:, c])
    s.close()
    return np.array(nodes, np.float64), np.array(mom), xp[:, 0]


def kirchhoff_from_energy(J, h=1e-6):
    """tau = J dpsi/dJ of psi(J) = K/2 (J - 1)^2, the derivative by FP64 central differences."""
    psi = lambda j: 0.5 * K_BULK * (j - 1.0) ** 2
    return J * (psi(J + h) - psi(J - h)) / (2.0 * h)


@pytest.mark.parametrize("J", [0.85, 1.15, 1.0])
def test_p2g_moment_is_minus_dt_vol_tau(J):
    dt, vol = 1e-4, DX ** 3 / 8
    nodes, mom, xp = one_particle_grid(J, dt, vol)
    assert len(nodes) == 27
    d = nodes * DX - xp[None, :]
    M = (mom[:, :, None] * d[:, None, :]).sum(0)
    tau = kirchhoff_from_energy(J)
    np.testing.assert_allclose(M, -dt * vol * tau * np.eye(3), rtol=1e-7, atol=1e-20)
    # the momentum itself sums to zero (the stress term moves none)
    np.testing.assert_allclose(mom.sum(0), 0.0, atol=1e-18)
    if J != 1.0:
        # the J factor is what separates tau = K J (J - 1) from K (J - 1): 15 % at these J
        assert abs(M[0, 0] / (-dt * vol * K_BULK * (J - 1.0)) - J) < 1e-6


@pytest.mark.parametrize("J,outward", [(0.9, True), (1.1, False)])
def test_compressed_block_pushes_out_expanded_pulls_in(J, outward):
    cells = np.stack(np.meshgrid(*[np.arange(12, 20)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(7)
    x = ((np.repeat(cells, 4, 0) + rng.random((cells.shape[0] * 4, 3))) * DX).T
    mats = [dict(model=O.MODEL_JFLUID, rho=1000.0, vol=DX ** 3 / 4, bulk=K_BULK)]
    s = O.Session(fluid_state(x, np.full(x.shape[1], J)), mats, DX, 1e-4, gravity=(0, 0, 0),
                  boundary=O.BOUNDARY_NONE, dtype=np.float64)
    s.substeps(1)
    out = s.state()
    s.close()
    c = x.mean(1, keepdims=True)
    r = x - c
    radial = (out["v"] * r).sum(0)
    face = np.abs(r).max(0) > 0.8 * np.abs(r).max()
    assert (radial[face].mean() > 0) == outward
    assert np.all(np.sign(radial[face]) == (1 if outward else -1))
    np.testing.assert_allclose(out["v"].sum(1), 0.0, atol=1e-12)
    # J = 1: no stress, nothing moves
    s1 = O.Session(fluid_state(x, np.ones(x.shape[1])), mats, DX, 1e-4, gravity=(0, 0, 0),
                   boundary=O.BOUNDARY_NONE, dtype=np.float64)
    s1.substeps(1)
    assert np.all(s1.state()["v"] == 0.0)
    s1.close()


def _column(J_profile):
    """A fluid column filling a 16-cell box with sticky walls of 3 cells: cells x, z in [2, 14),
    y in [2, 11), 8 jittered particles per cell.  Every free node (3 .. 13) has fluid on both
    sides, and the outer layer of particles reaches into the walls' nodes, which hold it."""
    n_cells, dx = 16, 1.0 / 16
    cells = np.stack(np.meshgrid(np.arange(2, 14), np.arange(2, 11), np.arange(2, 14), indexing="ij"),
                     -1).reshape(-1, 3)
    rng = np.random.default_rng(3)
    x = ((np.repeat(cells, 8, 0) + rng.random((cells.shape[0] * 8, 3))) * dx).T
    y_top = 11 * dx
    rho, K = 1000.0, 2.0e5
    J = J_profile(y_top - x[1], rho, K)
    mats = [dict(model=O.MODEL_JFLUID, rho=rho, vol=dx ** 3 / 8, bulk=K)]
    return fluid_state(x, J), mats, dx, n_cells, y_top


def test_hydrostatic_column_stays_at_rest():
    """Hydrostatic balance rho g = dp/dy with p = -K (J - 1): a column started on that profile
    stays near rest, and profiles for half or twice the pressure (what an EOS off by a factor 2
    in K, vol or the coupling would need) or no pressure at all move several times more."""
    g, dt, steps = 9.8, 2e-4, 50
    t = dt * steps
    rms = {}
    for name, scale in (("hydrostatic", 1.0), ("half", 0.5), ("double", 2.0), ("uncompressed", 0.0)):
        st, mats, dx, n_cells, y_top = _column(lambda d, rho, K: 1.0 - scale * rho * g * d / K)
        s = O.Session(st, mats, dx, dt, gravity=(0, -g, 0), pad=3, n_cells=n_cells, dtype=np.float64)
        s.substeps(steps)
        out = s.state()
        s.close()
        # particles below the free surface's cell
        deep = (st["x"][1] < y_top - 1.5 * dx)
        rms[name] = float(np.sqrt((out["v"][:, deep] ** 2).sum(0).mean()))
        if name == "hydrostatic":
            # J stays on the hydrostatic profile to a fraction of its range
            Jr = st["F"][0][deep]
            assert np.abs(out["F"][0][deep] - Jr).max() < 0.25 * (1.0 - Jr.min())
    # g t is the free-fall speed after the run
    assert rms["hydrostatic"] < 0.15 * g * t
    assert rms["half"] > 0.3 * g * t and rms["double"] > 0.3 * g * t
    assert rms["uncompressed"] > 0.5 * g * t
