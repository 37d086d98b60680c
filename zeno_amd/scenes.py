"""Host-side scene builders for the MPM workloads (particle initialisation).

The counterpart of the reference's ToZSParticles initialisation
(projects/CuLagrange/mpm/Generation.cpp:943-990): vol = dx^3 / ppc
(Generation.cu:25), m = rho * vol, v = v0, F = I (or J = 1 for the J-fluid),
C = 0.  Positions are jittered uniformly inside each cell from a counter RNG
(SplitMix64 -> 24-bit mantissa floats), so every scene is reproducible from its
seed on any host.

State convention (SoA, numpy): x, v (3, n); C, F (9, n) row-major; mat (n,)
int32 indices into the scene's material list.  J-fluid particles keep J in F[0].
"""
from __future__ import annotations

import math

import numpy as np

from .zmpm import MODEL_FCR, MODEL_JFLUID

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, counter: np.ndarray) -> np.ndarray:
    """SplitMix64 (Steele, Lea, Flood 2014) evaluated at seed + (counter+1) * golden gamma."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + (counter.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed: int, counter: np.ndarray) -> np.ndarray:
    """Uniform floats in [0, 1) with a 24-bit mantissa."""
    return (splitmix64(seed, counter) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def jitter_cells(cells: np.ndarray, ppc: int, dx: float, seed: int, counter0: int = 0) -> np.ndarray:
    """ppc jittered positions per cell; cells (k, 3) int -> positions (3, k*ppc) float32."""
    k = cells.shape[0]
    idx = np.arange(k * ppc * 3, dtype=np.uint64) + np.uint64(counter0)
    u = uniform24(seed, idx).reshape(k * ppc, 3)
    base = np.repeat(cells.astype(np.float64), ppc, axis=0)
    return ((base + u) * dx).astype(np.float32).T.copy()


def _state(x, v0, mats_idx, model_of):
    n = x.shape[1]
    v = np.zeros((3, n), np.float32)
    v[:] = np.asarray(v0, np.float32).reshape(3, -1) if np.ndim(v0) > 1 else np.asarray(v0, np.float32)[:, None]
    C = np.zeros((9, n), np.float32)
    F = np.zeros((9, n), np.float32)
    fluid = model_of[mats_idx] == MODEL_JFLUID
    F[0] = 1.0
    F[4] = np.where(fluid, 0.0, 1.0)
    F[8] = np.where(fluid, 0.0, 1.0)
    return {"x": x, "v": v, "C": C, "F": F, "mat": mats_idx.astype(np.int32)}


def concat_states(states):
    return {k: np.concatenate([s[k] for s in states], axis=-1) for k in states[0]}


def box_cells(lo, hi):
    """All integer cells in [lo, hi) (per axis), x-major order."""
    g = np.stack(np.meshgrid(*[np.arange(lo[k], hi[k]) for k in range(3)], indexing="ij"), axis=-1)
    return g.reshape(-1, 3)


def elastic_box(dx, lo_cell, hi_cell, ppc, seed, E, nu, rho=1000.0, v0=(0.0, 0.0, 0.0)):
    """A jittered fixed-corotated box; returns (state, materials)."""
    cells = box_cells(lo_cell, hi_cell)
    x = jitter_cells(cells, ppc, dx, seed)
    mats = [dict(model=MODEL_FCR, E=E, nu=nu, rho=rho, vol=dx ** 3 / ppc)]
    st = _state(x, v0, np.zeros(x.shape[1], np.int32), np.array([MODEL_FCR]))
    return st, mats


def config_plain_cube(ppc=2, seed=1):
    """Plain 3D MLS-MPM: cube of side 0.25 at the domain centre, dx = 1/128, fixed corotated
    E = 5e4, nu = 0.2, rho = 1e3.  ppc = 2 gives exactly 65,536 particles (cells [48, 80)^3)."""
    dx = 1.0 / 128
    st, mats = elastic_box(dx, (48, 48, 48), (80, 80, 80), ppc, seed, E=5e4, nu=0.2)
    return dict(name="plain_cube_65k", state=st, mats=mats, dx=dx, n_cells=128, dt=1e-4, steps=200,
                gravity=(0.0, -9.8, 0.0), pad=3)


def config_dam_break(total=4194304, seed=3):
    """Throughput: dx = 1/256; J-fluid dam column x in [3dx, 0.4], y in [3dx, 0.8] filled in z-slabs
    from z = 3dx at 8 ppc up to total - N_sphere particles (K = 1e4, rho = 1e3), plus a fixed-corotated
    sphere r = 0.08 at (0.7, 0.5, 0.5), E = 1e5, nu = 0.4, v0 = (0, -2, 0), 8 ppc.  dt = 5e-5."""
    dx = 1.0 / 256
    n = 256
    ppc = 8
    vol = dx ** 3 / ppc
    # sphere
    c = np.array([0.7, 0.5, 0.5])
    r = 0.08
    lo = np.floor((c - r) / dx).astype(int)
    hi = np.ceil((c + r) / dx).astype(int)
    scells = box_cells(lo, hi)
    xs = jitter_cells(scells, ppc, dx, seed, counter0=0)
    inside = ((xs.astype(np.float64) - c[:, None]) ** 2).sum(0) < r * r
    xs = xs[:, inside].copy()
    n_sphere = xs.shape[1]
    n_fluid = total - n_sphere
    # fluid column, z-slab order
    x_hi = int(math.floor(0.4 / dx))   # cells [3, 102): x <= 0.4
    y_hi = int(math.floor(0.8 / dx))   # cells [3, 204): y <= 0.8
    per_slab = (x_hi - 3) * (y_hi - 3) * ppc
    nslab = -(-n_fluid // per_slab)
    fcells = box_cells((0, 3, 3), (nslab, x_hi, y_hi))[:, [1, 2, 0]]  # (z, x, y) loop order -> (x, y, z)
    ncell_needed = -(-n_fluid // ppc)
    fcells = fcells[:ncell_needed]
    fcells[:, 2] += 3
    xf = jitter_cells(fcells, ppc, dx, seed, counter0=1 << 40)[:, :n_fluid].copy()
    mats = [dict(model=MODEL_JFLUID, rho=1000.0, vol=vol, bulk=1e4),
            dict(model=MODEL_FCR, E=1e5, nu=0.4, rho=1000.0, vol=vol)]
    model_of = np.array([MODEL_JFLUID, MODEL_FCR])
    sf = _state(xf, (0.0, 0.0, 0.0), np.zeros(n_fluid, np.int32), model_of)
    ss = _state(xs, (0.0, -2.0, 0.0), np.ones(n_sphere, np.int32), model_of)
    st = concat_states([sf, ss])
    return dict(name="dam_break_4M", state=st, mats=mats, dx=dx, n_cells=n, dt=5e-5, steps=100,
                gravity=(0.0, -9.8, 0.0), pad=3, n_fluid=n_fluid, n_solid=n_sphere)


def config_diff_block(seed=5):
    """Differentiable workload: 32 x 32 x 16-cell elastic block at dx = 1/128, 8 ppc (131,072
    particles), centred; E = 5e4, nu = 0.3; dt = 1e-4, 64 substeps.

    Initial velocity: a drift (0.3, -0.5, 0.2) m/s, a stretch along x and a squeeze
    along y (rates 30/s and -20/s about the centre) and a seeded per-particle
    perturbation of 0.02 m/s per axis.  Without the initial deformation the block
    falls with F = I and E, nu would not change its motion."""
    dx = 1.0 / 128
    st, mats = elastic_box(dx, (48, 48, 56), (80, 80, 72), 8, seed, E=5e4, nu=0.3)
    n = st["x"].shape[1]
    c = st["x"].astype(np.float64).mean(axis=1, keepdims=True)
    noise = uniform24(seed + 77, np.arange(3 * n, dtype=np.uint64)).reshape(3, n) * 2.0 - 1.0
    st["v"] = (np.array([0.3, -0.5, 0.2])[:, None] + np.array([30.0, -20.0, 0.0])[:, None] * (st["x"] - c)
               + 0.02 * noise).astype(np.float32)
    return dict(name="diff_block_131k", state=st, mats=mats, dx=dx, n_cells=128, dt=1e-4, steps=64,
                gravity=(0.0, -9.8, 0.0), pad=3)


def small_mixed_scene(n_cells=32, seed=11, v_solid=(0.3, -0.5, 0.2)):
    """A small scene with both models (tests): an elastic box and a fluid slab."""
    dx = 1.0 / n_cells
    a = box_cells((6, 10, 6), (12, 16, 12))
    b = box_cells((16, 4, 8), (26, 9, 20))
    xa = jitter_cells(a, 4, dx, seed)
    xb = jitter_cells(b, 4, dx, seed, counter0=1 << 32)
    mats = [dict(model=MODEL_FCR, E=2e4, nu=0.3, rho=1000.0, vol=dx ** 3 / 4),
            dict(model=MODEL_JFLUID, rho=1000.0, vol=dx ** 3 / 4, bulk=5e3)]
    model_of = np.array([MODEL_FCR, MODEL_JFLUID])
    sa = _state(xa, v_solid, np.zeros(xa.shape[1], np.int32), model_of)
    sb = _state(xb, (0.0, 0.0, 0.0), np.ones(xb.shape[1], np.int32), model_of)
    return dict(name="mixed_small", state=concat_states([sa, sb]), mats=mats, dx=dx, n_cells=n_cells,
                dt=2e-4, steps=20, gravity=(0.0, -9.8, 0.0), pad=3)


def random_box_scene(scene_id, seed, n_range=(2048, 65536), dt=1e-4):
    """One scene of the batched workload: a box of side U[0.1, 0.3] at the domain centre
    holding n = round(exp(U[ln n_lo, ln n_hi])) particles at uniform random positions,
    fixed corotated with E log-uniform in [1e4, 1e6], nu ~ U[0.1, 0.4], v0 ~ U[-1, 1]^3,
    rho = 1e3, vol = box volume / n, dx = 1/64 or 1/128 (fair coin).  All draws come from
    SplitMix64 at seed + scene_id, so every scene is reproducible on its own."""
    s = seed + scene_id
    u = uniform24(s, np.arange(8, dtype=np.uint64))
    n = int(round(math.exp(math.log(n_range[0]) + u[0] * (math.log(n_range[1]) - math.log(n_range[0])))))
    E = float(math.exp(math.log(1e4) + u[1] * (math.log(1e6) - math.log(1e4))))
    nu = float(0.1 + 0.3 * u[2])
    n_cells = 64 if u[3] < 0.5 else 128
    dx = 1.0 / n_cells
    side = float(0.1 + 0.2 * u[4])
    v0 = tuple(float(-1.0 + 2.0 * u[5 + k]) for k in range(3))
    p = uniform24(s, np.arange(8, 8 + 3 * n, dtype=np.uint64)).reshape(3, n)
    x = (0.5 - 0.5 * side + side * p).astype(np.float32)
    mats = [dict(model=MODEL_FCR, E=E, nu=nu, rho=1000.0, vol=side ** 3 / n)]
    st = _state(x, v0, np.zeros(n, np.int32), np.array([MODEL_FCR]))
    return dict(name=f"box_{scene_id}", state=st, mats=mats, dx=dx, n_cells=n_cells, dt=dt,
                gravity=(0.0, -9.8, 0.0), pad=3, E=E, nu=nu, side=side, v0=v0)


def config_batched(n_scenes=256, seed=4):
    """Batched workload: n_scenes independent random boxes (random_box_scene), 10 substeps per step."""
    return [random_box_scene(i, seed) for i in range(n_scenes)]


def merge_scenes(scenes):
    """Concatenate single-scene configs into one multi-scene context's inputs.

    Returns (scene descriptors, materials with their scene index, state with
    material indices offset into the merged table, per-scene particle ranges)."""
    descs, mats, states, ranges = [], [], [], []
    off = 0
    for i, sc in enumerate(scenes):
        descs.append(dict(dx=sc["dx"], n_cells=sc["n_cells"], dt=sc["dt"], gravity=sc["gravity"],
                          boundary=1, pad=sc["pad"]))
        st = dict(sc["state"])
        st["mat"] = st["mat"] + len(mats)
        mats.extend(dict(m, scene=i) for m in sc["mats"])
        n = st["x"].shape[1]
        ranges.append((off, off + n))
        off += n
        states.append(st)
    return descs, mats, concat_states(states), ranges
