"""The features added on top of the APIC substep (plasticity, level-set colliders, the affine
FLIP transfers) together in multi-scene contexts: HIP graph replays equal eager substeps bit for
bit (with eager work and synchronisation between replays), and each scene of a batch evolves
exactly as it does alone."""
import numpy as np
import pytest
import torch

from zeno_amd import scenes
from zeno_amd.zmpm import (BOUNDARY_STICKY_BOX, Context, StepParams, collider, MODEL_STVK_HENCKY,
                           PLASTIC_DRUCKER_PRAGER, PLASTIC_VON_MISES, TRANSFER_AFLIP, COLLIDER_SLIP,
                           COLLIDER_STICKY, SHAPE_SPHERE, SHAPE_HALFSPACE)

pytestmark = pytest.mark.gpu


def _dev(st):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for k, v in st.items()}


def _sand(scene_seed):
    sc = scenes.random_box_scene(scene_seed, 200, n_range=(1500, 4000))
    m = sc["mats"][0]
    plastic = PLASTIC_DRUCKER_PRAGER if scene_seed % 2 == 0 else PLASTIC_VON_MISES
    sc["mats"] = [dict(m, model=MODEL_STVK_HENCKY, plasticity=plastic, friction_angle=30.0,
                       yield_stress=0.05 * m["E"], flip_beta=0.5 * (scene_seed % 2))]
    return sc


def _colliders(batch):
    """A slip floor tilted a little in scene 0 and a moving sticky sphere in the last scene."""
    last = len(batch) - 1
    return [collider(type=COLLIDER_SLIP, shape=SHAPE_HALFSPACE, params=(0.0, 1.0, 0.0, 0.0),
                     translation=(0.5, 0.45, 0.5), ypr_degrees=(0.0, 8.0, 0.0), scene=0),
            collider(type=COLLIDER_STICKY, shape=SHAPE_SPHERE, params=(0.0, 0.0, 0.0, 0.04),
                     translation=(0.5, 0.5, 0.5), translation_rate=(0.3, 0.0, 0.0), scene=last)]


def _ctx(batch, transfer):
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    ctx = Context.scenes(descs, st["x"].shape[1])
    ctx.set_materials(mats)
    ctx.set_transfer(transfer)
    ctx.import_state(_dev(st))
    ctx.set_colliders(_colliders(batch))
    return ctx, ranges


@pytest.mark.parametrize("transfer", [0, TRANSFER_AFLIP])
def test_features_graph_replays_match_eager(transfer):
    batch = [_sand(i) for i in range(3)]
    eager, _ = _ctx(batch, transfer)
    graph, _ = _ctx(batch, transfer)
    graph.graph_capture(None, 4)
    for rep in range(3):
        eager.substeps(None, 4)
        graph.graph_launch()
        e = eager.export_state()
        g = graph.export_state()
        for k in ("x", "v", "C", "F"):
            assert torch.equal(e[k], g[k]), (rep, k)
    eager.close()
    graph.close()


def test_features_batch_equals_scenes_alone():
    """Every scene of a plastic, AFLIP batch with its colliders equals its own one-scene run."""
    batch = [_sand(i) for i in range(3)]
    ctx, ranges = _ctx(batch, TRANSFER_AFLIP)
    ctx.substeps(None, 6)
    out = ctx.export_state()
    cols = _colliders(batch)
    for s, sc in enumerate(batch):
        n = sc["state"]["x"].shape[1]
        one = Context(sc["dx"], sc["n_cells"], n)
        one.set_materials(sc["mats"])
        one.set_transfer(TRANSFER_AFLIP)
        one.import_state(_dev(sc["state"]))
        mine = []
        for c in cols:
            if c.scene == s:
                c2 = type(c).from_buffer_copy(c)
                c2.scene = 0
                mine.append(c2)
        one.set_colliders(mine)
        one.substeps(StepParams(sc["dt"], sc["gravity"], BOUNDARY_STICKY_BOX, sc["pad"]), 6)
        alone = one.export_state()
        a, b = ranges[s]
        for k in ("x", "v", "C", "F"):
            assert torch.equal(out[k][:, a:b], alone[k]), (s, k)
        one.close()
    ctx.close()


def test_per_particle_lame(oracle):
    """Per-particle mu, lam (Transfer.cu:861-864): the material's own values change nothing, bit for
    bit, in the substeps and in the adjoint; other values match the oracle and travel with their
    particles; FLIP (no override in p2g_flip) ignores them."""
    import numpy as np
    from zeno_amd import scenes
    from zeno_amd.zmpm import Context, StepParams, TRANSFER_FLIP, MODEL_JFLUID
    dev = "cuda:0"
    cfg = scenes.small_mixed_scene()
    st = cfg["state"]
    n = st["x"].shape[1]
    mu_m = np.zeros(n, np.float32)
    lam_m = np.zeros(n, np.float32)
    for k, m in enumerate(cfg["mats"]):
        if m["model"] != MODEL_JFLUID:
            mu, lam = oracle.lame_f32(m["E"], m["nu"])
            mu_m[st["mat"] == k], lam_m[st["mat"] == k] = mu, lam
    p = StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])

    def ctx_for(extra, scheme=None):
        c = Context(cfg["dx"], cfg["n_cells"], n)
        c.set_materials(cfg["mats"])
        if scheme is not None:
            c.set_transfer(scheme)
        c.import_state({k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in dict(st, **extra).items()})
        return c

    same = dict(mu=mu_m, lam=lam_m)
    a, b = ctx_for({}), ctx_for(same)
    # the adjoint of one substep, same output adjoints
    rng = np.random.default_rng(5)
    adj0 = {k: torch.from_numpy(rng.normal(size=s).astype(np.float32)).to(dev)
            for k, s in (("x", (3, n)), ("v", (3, n)), ("C", (9, n)), ("F", (9, n)))}
    outs = []
    for c in (a, b):
        adj = {k: v.clone() for k, v in adj0.items()}
        adj["mu"] = torch.zeros(n, device=dev)
        adj["lam"] = torch.zeros(n, device=dev)
        c.adjoint_substep(p, adj)
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in adj.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    a.substeps(p, 10)
    b.substeps(p, 10)
    ea, eb = a.export_state(dev), b.export_state(dev)
    for k in ea:
        assert torch.equal(ea[k], eb[k]), k
    a.close()
    b.close()
    # other values: parity with the oracle, carried through the block moves
    scale = rng.uniform(0.5, 1.5, n).astype(np.float32)
    other = dict(mu=mu_m * scale, lam=lam_m * scale[::-1].copy())
    c = ctx_for(other)
    c.substeps(p, 20)
    got = {k: v.cpu().numpy() for k, v in c.export_state(dev, aux=True).items()}
    c.close()
    o = oracle.Session(st, cfg["mats"], cfg["dx"], cfg["dt"])
    o.set_lame(other["mu"], other["lam"])
    o.substeps(20)
    exp = o.state()
    assert np.abs(got["x"] - exp["x"]).max() / cfg["dx"] <= 1e-3
    assert np.abs(got["v"] - exp["v"]).max() <= 1e-3 * max(np.abs(exp["v"]).max(), 1e-3)
    assert np.array_equal(got["mu"], other["mu"]) and np.array_equal(got["lam"], other["lam"])
    d = ctx_for({}, TRANSFER_FLIP)
    e = ctx_for(other, TRANSFER_FLIP)
    d.substeps(p, 5)
    e.substeps(p, 5)
    ed, ee = d.export_state(dev), e.export_state(dev)
    for k in ed:
        assert torch.equal(ed[k], ee[k]), k
    d.close()
    e.close()
