"""The fused G2P / P2G pass against the node chain and the six-launch substep chain.

With zmpm_set_fused (or the autotuner's choice), zmpm_substeps runs one kernel
per block of the new partition that does this substep's G2P and the next
substep's stress and P2G.  The pass sums a block's P2G contributions one chunk
of 384 particles at a time (registers cleared per chunk); while it is on, every
P2G of the context groups its sums the same way.  Bars:
- fused substeps = the node chain of the same context (partition, P2G, grid
  update, boundary, G2P), bit for bit, and the same substeps split over several
  calls (3 + 4 + 5 against 12: v and C are stored only by a call's last substep,
  and the tiles of the next P2G carry over between calls), bit for bit;
- fused = the six-launch chain (registers carried across chunks) bit for bit
  when no block holds more than 384 particles, otherwise within FP32 rounding of
  the per-chunk grouping: |dx| <= 1e-5 cells, |dv| <= 1e-5 |v|inf after 12 substeps;
- a dt change and a re-import between calls (the tiles are recomputed).
The scenes exercise every path of the pass: stayers (grid velocities staged in
LDS) and arrivals (gathered from HBM: many when the solid moves half a cell per
substep), blocks of three chunks (16 particles per cell), fixed corotated,
StVK-Hencky with Drucker-Prager, von Mises and Cam-clay return mapping (logJp
carried), the ARAP fiber, per-particle Lame parameters, mass and volume,
J-fluid, and several scenes in one context.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from zeno_amd import scenes  # noqa: E402
from zeno_amd.zmpm import (ANISO_ARAP, BOUNDARY_STICKY_BOX, MODEL_FCR, MODEL_JFLUID,  # noqa: E402
                           MODEL_STVK_HENCKY, PLASTIC_CAM_CLAY, PLASTIC_DRUCKER_PRAGER, PLASTIC_VON_MISES,
                           Context, StepParams)

DEV = "cuda:0"


def to_dev(state):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in state.items()}


def boxes_scene(mats, seed=5, ppc=4, n_cells=32, v=None, extra=None):
    """One jittered 6^3-cell box per material, side by side; v: per-box initial velocity."""
    dx = 1.0 / n_cells
    parts = []
    for k in range(len(mats)):
        lo = (4 + 7 * (k % 3), 8 + 8 * (k // 3), 10)
        x = scenes.jitter_cells(scenes.box_cells(lo, tuple(c + 6 for c in lo)), ppc, dx, seed + k)
        parts.append((x, k))
    model_of = np.array([m["model"] for m in mats])
    sts = []
    for x, k in parts:
        vk = (v[k] if v is not None else (0.3, -0.5 + 0.2 * k, 0.1))
        sts.append(scenes._state(x, vk, np.full(x.shape[1], k, np.int32), model_of))
    st = scenes.concat_states(sts)
    n = st["x"].shape[1]
    rng = np.random.default_rng(seed)
    c = st["x"].mean(axis=1, keepdims=True)
    st["v"] = (st["v"] + np.array([20.0, -15.0, 5.0])[:, None] * (st["x"] - c)
               + 0.05 * rng.standard_normal((3, n))).astype(np.float32)
    if extra:
        st.update(extra(st, rng))
    return dict(state=st, mats=mats, dx=dx, n_cells=n_cells, dt=2e-4, gravity=(0.0, -9.8, 0.0), pad=3)


def make_ctx(cfg, scenes_desc=None):
    st = cfg["state"]
    n = st["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], n) if scenes_desc is None else Context.scenes(scenes_desc, n)
    ctx.set_materials(cfg["mats"])
    ctx.import_state(to_dev(st))
    return ctx


def finish(ctx):
    out = {k: v.cpu().numpy() for k, v in ctx.export_state(DEV, aux=True).items()}
    keys, data = ctx.grid()
    ctx.close()
    return out, keys, data


def run(cfg, splits, fused, dts=None, reimport_after=None, scenes_desc=None):
    st = cfg["state"]
    ctx = make_ctx(cfg, scenes_desc)
    ctx.set_fused(fused)
    assert ctx.fused() == fused
    for k, steps in enumerate(splits):
        if scenes_desc is None:
            p = StepParams(dts[k] if dts else cfg["dt"], cfg["gravity"], BOUNDARY_STICKY_BOX, cfg["pad"])
        else:
            p = None
        ctx.substeps(p, steps)
        if reimport_after is not None and k == reimport_after:
            mid = ctx.export_state(DEV, aux=True)
            mid["mat"] = torch.from_numpy(np.ascontiguousarray(st["mat"])).to(DEV)
            ctx.import_state(mid)
    return finish(ctx)


def node_chain(cfg, steps):
    """The stage calls of a fused context (its P2G groups sums per chunk as the pass does)."""
    ctx = make_ctx(cfg)
    ctx.set_fused(True)
    p = StepParams(cfg["dt"], cfg["gravity"], BOUNDARY_STICKY_BOX, cfg["pad"])
    for _ in range(steps):
        ctx.partition()
        ctx.grid_clear()
        ctx.p2g(p.dt)
        ctx.grid_update(p.dt, p.gravity)
        ctx.boundary(p.boundary, p.pad)
        ctx.g2p(p.dt)
        ctx.return_mapping()
    return finish(ctx)


def assert_same(a, b):
    for k in a[0]:
        assert np.array_equal(a[0][k], b[0][k], equal_nan=True), k
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def assert_close(a, b, dx):
    assert np.abs(a[0]["x"] - b[0]["x"]).max() <= 1e-5 * dx
    assert np.abs(a[0]["v"] - b[0]["v"]).max() <= 1e-5 * max(np.abs(b[0]["v"]).max(), 1e-3)


def check(cfg, multi_chunk=False, scenes_desc=None):
    fused = run(cfg, [12], True, scenes_desc=scenes_desc)
    assert_same(run(cfg, [3, 4, 5], True, scenes_desc=scenes_desc), fused)
    if scenes_desc is None:
        assert_same(node_chain(cfg, 12), fused)
    plain = run(cfg, [12], False, scenes_desc=scenes_desc)
    if multi_chunk:
        assert_close(plain, fused, cfg.get("dx", 1.0 / 128))
    else:
        assert_same(plain, fused)
    return fused


def test_mixed_scene_with_fast_solid():
    """Fixed corotated + J-fluid; the solid moves about half a cell per substep (many arrivals)."""
    check(scenes.small_mixed_scene(v_solid=(40.0, -60.0, 25.0)))


def test_multi_chunk_blocks():
    """16 particles per cell: every interior block holds 1,024 particles, three P2G chunks."""
    mats = [dict(model=MODEL_FCR, E=3e4, nu=0.3, rho=1000.0, vol=(1 / 32) ** 3 / 16),
            dict(model=MODEL_JFLUID, rho=1000.0, vol=(1 / 32) ** 3 / 16, bulk=5e3)]
    check(boxes_scene(mats, ppc=16), multi_chunk=True)


def test_plasticity_and_stvk():
    mats = [dict(model=MODEL_STVK_HENCKY, E=2e4, nu=0.3, rho=1500.0, vol=1e-6, plasticity=PLASTIC_DRUCKER_PRAGER,
                 friction_angle=30.0),
            dict(model=MODEL_STVK_HENCKY, E=2e4, nu=0.3, rho=1500.0, vol=1e-6, plasticity=PLASTIC_VON_MISES,
                 yield_stress=300.0),
            dict(model=MODEL_STVK_HENCKY, E=2e4, nu=0.3, rho=1500.0, vol=1e-6),
            dict(model=MODEL_FCR, E=2e4, nu=0.3, rho=1500.0, vol=1e-6),
            dict(model=MODEL_STVK_HENCKY, E=4e4, nu=0.25, rho=1000.0, vol=1e-6, plasticity=PLASTIC_CAM_CLAY,
                 friction_angle=40.0)]
    check(boxes_scene(mats))


def test_arap_and_per_particle_parameters():
    """The ARAP fiber, per-particle mu, lam, m and vol (all carried by the pass)."""
    mats = [dict(model=MODEL_FCR, E=3e4, nu=0.3, rho=1000.0, vol=1e-6, aniso=ANISO_ARAP, aniso_strength=5.0),
            dict(model=MODEL_STVK_HENCKY, E=3e4, nu=0.3, rho=1000.0, vol=1e-6, aniso=ANISO_ARAP),
            dict(model=MODEL_JFLUID, rho=1000.0, vol=1e-6, bulk=4e3)]

    def extra(st, rng):
        n = st["x"].shape[1]
        a = rng.standard_normal((3, n))
        a /= np.linalg.norm(a, axis=0, keepdims=True)
        vol = (1e-6 * rng.uniform(0.8, 1.2, n)).astype(np.float32)
        return dict(a=a.astype(np.float32), mu=rng.uniform(8e3, 1.5e4, n).astype(np.float32),
                    lam=rng.uniform(1e4, 2e4, n).astype(np.float32), vol=vol,
                    m=(1000.0 * vol * rng.uniform(0.9, 1.1, n)).astype(np.float32))
    check(boxes_scene(mats, extra=extra))


def test_dt_change_and_reimport_between_calls():
    cfg = scenes.small_mixed_scene(v_solid=(10.0, -20.0, 5.0))
    a = run(cfg, [4, 4, 4], True, dts=[2e-4, 1e-4, 2e-4])
    b = run(cfg, [4, 4, 4], False, dts=[2e-4, 1e-4, 2e-4])
    assert_same(a, b)
    a = run(cfg, [5, 7], True, reimport_after=0)
    b = run(cfg, [5, 7], False, reimport_after=0)
    assert_same(a, b)


def test_several_scenes_in_one_context():
    """Each scene of a fused batch evolves as it does alone in a fused context."""
    batch = [scenes.random_box_scene(i, 300, n_range=(500, 5000)) for i in range(4)]
    batch.insert(1, scenes.small_mixed_scene(v_solid=(20.0, -30.0, 10.0)))
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    cfg = dict(state=st, mats=mats)
    fused = check(cfg, multi_chunk=True, scenes_desc=descs)
    for i, sc in enumerate(batch):
        alone = run(sc, [12], True)
        a, b = ranges[i]
        for k in ("x", "v", "C", "F"):
            assert np.array_equal(fused[0][k][:, a:b], alone[0][k]), (i, k)
