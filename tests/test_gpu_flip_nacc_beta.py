"""GPU parity of the non-affine FLIP transfer (p2g_flip Transfer.cu:892-933, the FLIP G2P
Transfer.cu:1215-1250), the Cam-clay return mapping ("nacc", Generation.cu:77-82 with
ZSReturnMapping MPMPipeline.cu:958-966) and ComputeParticleBeta (Boundary.cu:284-313)
against the oracle, and the particles' beta / logJp channels.

Bars:
  FLIP, 20 substeps, elastic box + J-fluid slab ........ |dx| / dx <= 1e-3 cells, |dv| <= 1e-3 |v|inf;
                                                        C kept bit for bit; differs from AFLIP
  FLIP fused vs node chain ............................. bit-exact
  Cam-clay of one F (FP32 vs FP64 oracle) .............. |dF| <= 2e-5 |F| + 1e-6, |d logJp| <= 1e-5
  Cam-clay substeps (snow block on the floor) .......... |dx| / dx <= 1e-3 cells, |dF| <= 1e-3,
                                                        |d logJp| <= 1e-4
  ComputeParticleBeta .................................. beta equal (away from phi = 0, J = jpc and
                                                        a zero normal speed, where FP32 rounding decides)
  beta / logJp import, export, scatter .................. bit-exact
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from zeno_amd import scenes  # noqa: E402
from zeno_amd.zmpm import (Context, StepParams, collider, MODEL_FCR, MODEL_STVK_HENCKY,  # noqa: E402
                           PLASTIC_CAM_CLAY, TRANSFER_AFLIP, TRANSFER_APIC, TRANSFER_FLIP, COLLIDER_STICKY,
                           SHAPE_HALFSPACE, SHAPE_SPHERE)

DEV = "cuda:0"


def to_dev(state):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in state.items()}


def flip_scene():
    """An elastic box and a J-fluid slab with per-particle velocity noise (the FLIP part)."""
    cfg = scenes.small_mixed_scene(v_solid=(2.0, -3.0, 1.0))
    st = dict(cfg["state"])
    rng = np.random.default_rng(17)
    st["v"] = (st["v"] + 0.5 * rng.normal(size=st["v"].shape)).astype(np.float32)
    st["C"] = (0.3 * rng.normal(size=st["C"].shape)).astype(np.float32)  # FLIP must keep it
    cfg["state"] = st
    return cfg


def run(cfg, steps, scheme, fused=True, chain_extra=None, aux=False):
    n = cfg["state"]["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    ctx.set_transfer(scheme)
    ctx.import_state(to_dev(cfg["state"]))
    p = StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])
    if fused:
        ctx.substeps(p, steps)
    else:
        for _ in range(steps):
            ctx.partition()
            ctx.grid_clear()
            ctx.p2g(p.dt)
            ctx.grid_update(p.dt, p.gravity)
            ctx.boundary(p.boundary, p.pad)
            ctx.g2p(p.dt)
            if chain_extra:
                chain_extra(ctx)
    out = {k: v.cpu().numpy() for k, v in ctx.export_state(DEV, aux=aux).items()}
    ctx.close()
    return out


def test_flip_substeps_match_oracle(oracle):
    cfg = flip_scene()
    steps = 20
    a = run(cfg, steps, TRANSFER_FLIP)
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    o.set_transfer(oracle.TRANSFER_FLIP)
    o.substeps(steps)
    b = o.state()
    dx = cfg["dx"]
    assert np.abs(a["x"] - b["x"]).max() / dx <= 1e-3
    assert np.abs(a["v"] - b["v"]).max() <= 1e-3 * max(np.abs(b["v"]).max(), 1e-3)
    assert np.array_equal(a["C"], cfg["state"]["C"])  # FLIP writes no C
    c = run(cfg, steps, TRANSFER_AFLIP)
    assert np.abs(a["v"] - c["v"]).max() > 1e-3


def test_flip_fused_equals_node_chain():
    cfg = flip_scene()
    a = run(cfg, 10, TRANSFER_FLIP, fused=True)
    b = run(cfg, 10, TRANSFER_FLIP, fused=False)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ---------------------------------------------------------------- Cam-clay
NACC = dict(model=MODEL_STVK_HENCKY, E=2e4, nu=0.3, rho=400.0, vol=1e-6, plasticity=PLASTIC_CAM_CLAY,
            friction_angle=35.0, nacc_beta=2.0, nacc_xi=1.0)


def rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_cam_clay_return_mapping_matches_oracle(oracle):
    rng = np.random.default_rng(41)
    n = 3000
    Fs = []
    for _ in range(n):
        e = rng.normal(size=3)
        e = 0.05 * (e - e.mean()) + rng.uniform(-0.08, 0.05) / 3.0
        Fs.append(rot(rng) @ np.diag(np.exp(e)) @ rot(rng))
    F = np.stack(Fs).reshape(n, 9).T.astype(np.float32)
    logJp = rng.uniform(-0.1, 0.0, n).astype(np.float32)
    x = rng.uniform(0.3, 0.7, size=(3, n)).astype(np.float32)
    st = dict(x=x, v=np.zeros((3, n), np.float32), C=np.zeros((9, n), np.float32), F=F,
              mat=np.zeros(n, np.int32), logJp=logJp)
    ctx = Context(1.0 / 64, 64, n)
    ctx.set_materials([NACC])
    ctx.import_state(to_dev(st))
    ctx.return_mapping()
    out = {k: v.cpu().numpy() for k, v in ctx.export_state(DEV, aux=True).items()}
    ctx.close()
    mu, lam = oracle.lame_f32(2e4, 0.3)
    moved = tips = 0
    for i in range(n):
        exp, lj, changed = oracle.project_cam_clay(F[:, i].astype(np.float64), mu, lam, 35.0, 2.0, 1.0,
                                                   float(logJp[i]))
        moved += changed
        tips += lj != float(logJp[i])
        got = out["F"][:, i].astype(np.float64).reshape(3, 3)
        assert np.abs(got - exp).max() <= 2e-5 * np.abs(exp).max() + 1e-6, (i, got, exp)
        assert abs(out["logJp"][i] - lj) <= 1e-5, (i, out["logJp"][i], lj)
    assert moved > n // 4 and tips > 20  # both the surface and the tips are exercised


def snow_scene():
    """A Cam-clay snow block dropped onto the sticky floor (dx = 1/32, 4 particles per cell)."""
    dx = 1.0 / 32
    x = scenes.jitter_cells(scenes.box_cells((8, 3, 8), (16, 10, 16)), 4, dx, 51)
    n = x.shape[1]
    mats = [dict(NACC, vol=dx ** 3 / 4)]
    st = scenes._state(x, (0.0, -2.0, 0.0), np.zeros(n, np.int32), np.array([MODEL_STVK_HENCKY]))
    return dict(state=st, mats=mats, dx=dx, n_cells=32, dt=2e-4, gravity=(0.0, -9.8, 0.0), pad=3)


def test_cam_clay_substeps_match_oracle(oracle):
    cfg = snow_scene()
    steps = 30
    a = run(cfg, steps, TRANSFER_APIC, aux=True)
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    o.substeps(steps)
    b = o.state()
    _, lj = o.aux()
    dx = cfg["dx"]
    assert np.abs(a["x"] - b["x"]).max() / dx <= 1e-3
    assert np.abs(a["v"] - b["v"]).max() <= 1e-3 * max(np.abs(b["v"]).max(), 1e-3)
    assert np.abs(a["F"] - b["F"]).max() <= 1e-3
    assert np.abs(a["logJp"] - lj).max() <= 1e-4
    assert (lj != np.float32(-0.04)).sum() > 10  # the snow compacted


def test_cam_clay_fused_equals_node_chain():
    cfg = snow_scene()
    a = run(cfg, 12, TRANSFER_APIC, aux=True)
    b = run(cfg, 12, TRANSFER_APIC, fused=False, chain_extra=lambda c: c.return_mapping(), aux=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ---------------------------------------------------------------- ComputeParticleBeta
def beta_scene(n=4000, seed=61):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(3, n)).astype(np.float32)
    v = rng.normal(scale=0.5, size=(3, n)).astype(np.float32)
    F = np.zeros((9, n), np.float32)
    for i in range(n):
        F[:, i] = (np.eye(3) + rng.normal(scale=0.03, size=(3, 3))).reshape(9)
    st = dict(x=x, v=v, C=np.zeros((9, n), np.float32), F=F, mat=np.zeros(n, np.int32))
    mats = [dict(model=MODEL_FCR, E=1e4, nu=0.3, rho=1000.0, vol=1e-6, flip_beta=0.25)]
    cols = [dict(type=COLLIDER_STICKY, shape=SHAPE_SPHERE, params=(0.5, 0.5, 0.5, 0.12), translation_rate=(0.3, 0, 0)),
            dict(type=COLLIDER_STICKY, shape=SHAPE_HALFSPACE, params=(0.0, 1.0, 0.0, 0.42))]
    return st, mats, cols


def borderline(st, cols, dt, jpc):
    """Particles whose beta FP32 rounding may decide: |phi|, the normal speed or J - jpc tiny."""
    x = st["x"].astype(np.float64) + dt * st["v"].astype(np.float64)
    v = st["v"].astype(np.float64)
    J = np.linalg.det(st["F"].T.reshape(-1, 3, 3).astype(np.float64))
    bad = np.abs(J - jpc) < 1e-5
    d = x - np.array(cols[0]["params"][:3])[:, None]
    r = np.linalg.norm(d, axis=0)
    bad |= np.abs(r - cols[0]["params"][3]) < 1e-5
    bad |= np.abs((((v - np.array(cols[0]["translation_rate"])[:, None]) * d).sum(0)) / np.maximum(r, 1e-12)) < 1e-5
    bad |= np.abs(x[1] - cols[1]["params"][3]) < 1e-5
    bad |= np.abs(v[1]) < 1e-5
    return bad


def test_particle_beta_matches_oracle(oracle):
    st, mats, cols = beta_scene()
    dt, jpc = 1e-3, 1.0
    n = st["x"].shape[1]
    ctx = Context(1.0 / 64, 64, n)
    ctx.set_materials(mats)
    ctx.set_colliders([collider(**c) for c in cols])
    ctx.import_state(to_dev(st))
    ctx.compute_particle_beta(StepParams(dt), jpc)
    got = ctx.export_state(DEV, aux=True)["beta"].cpu().numpy()
    ctx.close()
    o = oracle.Session(st, mats, 1.0 / 64, dt)
    o.set_colliders(cols)
    o.compute_particle_beta(dt, jpc)
    exp, _ = o.aux()
    ok = ~borderline(st, cols, dt, jpc)
    assert ok.sum() > n - 20
    np.testing.assert_array_equal(got[ok], exp[ok])
    assert (exp == 0.0).sum() > 100 and (exp == 0.5).sum() > 100


def test_asflip_with_particle_beta_matches_oracle(oracle):
    """ComputeParticleBeta, then ASFLIP substeps reading each particle's beta."""
    cfg = flip_scene()
    cfg["mats"] = [dict(m, flip_beta=0.0) for m in cfg["mats"]]
    st = cfg["state"]
    n = st["x"].shape[1]
    beta = np.where(np.arange(n) % 3 == 0, 0.5, 0.0).astype(np.float32)
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    ctx.set_transfer(TRANSFER_AFLIP)
    ctx.import_state(to_dev(dict(st, beta=beta)))
    ctx.substeps(StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"]), 15)
    a = {k: v.cpu().numpy() for k, v in ctx.export_state(DEV, aux=True).items()}
    ctx.close()
    o = oracle.Session(st, cfg["mats"], cfg["dx"], cfg["dt"])
    o.set_transfer(oracle.TRANSFER_AFLIP)
    o.set_aux(beta=beta)
    o.substeps(15)
    b = o.state()
    assert np.array_equal(a["beta"], beta)  # carried with its particle through the block moves
    assert np.abs(a["x"] - b["x"]).max() / cfg["dx"] <= 1e-3
    c = run(cfg, 15, TRANSFER_AFLIP)  # beta 0 everywhere: a different trajectory
    assert np.abs(a["x"] - c["x"]).max() > 0.0


def test_aux_channels_import_export_and_defaults():
    cfg = snow_scene()
    st = cfg["state"]
    n = st["x"].shape[1]
    rng = np.random.default_rng(3)
    lj = rng.uniform(-3.0, -2.0, n).astype(np.float32)  # hardened: no projection in 5 substeps
    beta = rng.uniform(0, 1, n).astype(np.float32)
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    ctx.import_state(to_dev(dict(st, logJp=lj, beta=beta)))
    out = ctx.export_state(DEV, aux=True)
    assert np.array_equal(out["logJp"].cpu().numpy(), lj)
    # APIC does not carry beta: it reads as the material's flip_beta
    assert np.all(out["beta"].cpu().numpy() == np.float32(cfg["mats"][0].get("flip_beta", 0.0)))
    ctx.substeps(StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"]), 5)
    assert np.array_equal(ctx.export_state(DEV, aux=True)["logJp"].cpu().numpy(), lj)
    ctx.close()
    # defaults: logJp -0.04 (Generation.cpp:987), beta from the material
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials([dict(cfg["mats"][0], flip_beta=0.3)])
    ctx.set_transfer(TRANSFER_FLIP)
    ctx.import_state(to_dev(st))
    out = ctx.export_state(DEV, aux=True)
    assert np.all(out["logJp"].cpu().numpy() == np.float32(-0.04))
    assert np.all(out["beta"].cpu().numpy() == np.float32(0.3))
    ctx.close()


def test_particle_beta_default_jpc_per_material():
    """No jp_critical: the reference's JpcDefault, 1 for every material, Cam-clay included
    (Boundary.cu:389-391 tests a variant that never holds Cam-clay); an explicit 1.1 applies to all."""
    from zeno_amd.zmpm import MODEL_STVK_HENCKY, PLASTIC_CAM_CLAY
    rng = np.random.default_rng(4)
    n = 256
    x = rng.uniform(0.4, 0.6, size=(3, n)).astype(np.float32)
    F = np.zeros((9, n), np.float32)
    F[[0, 4, 8]] = 1.0
    F[0] = np.where(np.arange(n) % 4 < 2, 1.05, 1.15)
    mat = (np.arange(n) % 2).astype(np.int32)
    mats = [dict(model=MODEL_STVK_HENCKY, E=1e4, nu=0.3, rho=1000.0, vol=1.0 / 64 ** 3 / 8),
            dict(model=MODEL_STVK_HENCKY, E=1e4, nu=0.3, rho=1000.0, vol=1.0 / 64 ** 3 / 8,
                 plasticity=PLASTIC_CAM_CLAY)]
    st = dict(x=x, v=np.zeros((3, n), np.float32), C=np.zeros((9, n), np.float32), F=F, mat=mat)
    ctx = Context(1.0 / 64, 64, n)
    ctx.set_materials(mats)
    ctx.import_state(to_dev(st))
    ctx.compute_particle_beta(StepParams(1e-3))
    got = ctx.export_state(DEV, aux=True)["beta"].cpu().numpy()
    np.testing.assert_array_equal(got, np.full(n, 0.5, np.float32))
    ctx.compute_particle_beta(StepParams(1e-3), 1.1)
    got = ctx.export_state(DEV, aux=True)["beta"].cpu().numpy()
    exp = np.where(F[0] < 1.1, 0.0, 0.5).astype(np.float32)
    np.testing.assert_array_equal(got, exp)
    ctx.close()
