"""Particle blocks cut into parts: a block of more than 1,024 particles is several P2G and G2P
workgroups (the first part is the block itself; each extra part has its own tile, which the
grid kernel adds after the blocks' first parts, and its own move counts, its ranks offset by
the block's earlier parts).

Scenes of 27 particles per cell (1,728 per interior 4^3 block: two parts) and mixed
fixed-corotated / J-fluid blocks.  Bars:
  substeps (parts) vs the node chain (full re-bin each substep, the same parts): bit-exact
  repeat: bit-exact; graph replay vs eager substeps: bit-exact
  the dense scene inside a batch of scenes vs alone: bit-exact
  30 substeps vs the FP32 oracle: |dx| / dx <= 1e-3 cells, |dv| <= 1e-3 |v|inf (APIC and AFLIP)
  switching the fused pass (one part per block) on and off between calls, and a fused graph
  launched after unfused substeps: vs the oracle as above
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from zeno_amd import scenes  # noqa: E402
from zeno_amd.zmpm import (MODEL_FCR, MODEL_JFLUID, TRANSFER_AFLIP, Context, StepParams)  # noqa: E402

DEV = "cuda:0"


def dense_scene(seed=21, ppc=27):
    """A fixed-corotated box and a J-fluid slab at 27 particles per cell, dx = 1/32."""
    dx = 1.0 / 32
    xa = scenes.jitter_cells(scenes.box_cells((8, 10, 8), (16, 18, 16)), ppc, dx, seed)
    xb = scenes.jitter_cells(scenes.box_cells((18, 4, 6), (26, 9, 18)), ppc, dx, seed, counter0=1 << 32)
    mats = [dict(model=MODEL_FCR, E=4e4, nu=0.3, rho=1000.0, vol=dx ** 3 / ppc, flip_beta=0.5),
            dict(model=MODEL_JFLUID, rho=1000.0, vol=dx ** 3 / ppc, bulk=4e3)]
    model_of = np.array([MODEL_FCR, MODEL_JFLUID])
    sa = scenes._state(xa, (0.4, -0.6, 0.2), np.zeros(xa.shape[1], np.int32), model_of)
    sb = scenes._state(xb, (-0.3, -0.2, 0.1), np.ones(xb.shape[1], np.int32), model_of)
    st = scenes.concat_states([sa, sb])
    rng = np.random.default_rng(seed)
    c = st["x"].mean(axis=1, keepdims=True)
    st["v"] = (st["v"] + np.array([12.0, -9.0, 4.0])[:, None] * (st["x"] - c)
               + 0.05 * rng.standard_normal(st["x"].shape)).astype(np.float32)
    return dict(state=st, mats=mats, dx=dx, n_cells=32, dt=1e-4, gravity=(0.0, -9.8, 0.0), pad=3)


def make_ctx(cfg, transfer=None):
    n = cfg["state"]["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    if transfer is not None:
        ctx.set_transfer(transfer)
    ctx.import_state({k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in cfg["state"].items()})
    return ctx


def params(cfg):
    return StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])


def export(ctx):
    out = {k: v.cpu().numpy() for k, v in ctx.export_state(DEV).items()}
    ctx.sync()
    return out


def assert_same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def assert_oracle(a, cfg, steps, oracle, transfer=None):
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    if transfer is not None:
        o.set_transfer(transfer)
    o.substeps(steps)
    b = o.state()
    assert np.abs(a["x"] - b["x"]).max() / cfg["dx"] <= 1e-3
    assert np.abs(a["v"] - b["v"]).max() <= 1e-3 * max(np.abs(b["v"]).max(), 1e-3)


def test_blocks_are_split():
    """The parts of the current partition: ceil(count / 1024) per block of the current positions."""
    cfg = dense_scene()
    ctx = make_ctx(cfg)
    ctx.substeps(params(cfg), 2)
    nb, nparts = ctx.num_particle_blocks(), ctx.num_block_parts()
    x = export(ctx)["x"]
    b = np.floor(x * np.float32(32) - np.float32(0.5)).astype(np.int64) >> 2
    _, counts = np.unique((b[0] * 64 + b[1]) * 64 + b[2], return_counts=True)
    print(f"particle blocks {nb}, parts {nparts}, largest block {counts.max()}", flush=True)
    assert nb == counts.size
    assert nparts == int(((counts + 1023) // 1024).sum()) > nb
    ctx.set_fused(True)
    ctx.substeps(params(cfg), 2)
    assert ctx.num_block_parts() == ctx.num_particle_blocks()
    ctx.close()


def test_parts_substeps_equal_node_chain_and_repeat():
    cfg = dense_scene()
    p = params(cfg)
    runs = []
    for _ in range(2):
        ctx = make_ctx(cfg)
        ctx.substeps(p, 10)
        runs.append(export(ctx))
        ctx.close()
    assert_same(runs[0], runs[1])
    ctx = make_ctx(cfg)
    for _ in range(10):
        ctx.partition()
        ctx.grid_clear()
        ctx.p2g(p.dt)
        ctx.grid_update(p.dt, p.gravity)
        ctx.boundary(p.boundary, p.pad)
        ctx.g2p(p.dt)
    chain = export(ctx)
    ctx.close()
    assert_same(runs[0], chain)


def test_parts_graph_equals_eager():
    cfg = dense_scene()
    p = params(cfg)
    ctx = make_ctx(cfg)
    ctx.substeps(p, 2)
    ctx.graph_capture(p, 4)
    for _ in range(3):
        ctx.graph_launch()
    g = export(ctx)
    ctx.close()
    ctx = make_ctx(cfg)
    ctx.substeps(p, 14)
    e = export(ctx)
    ctx.close()
    assert_same(g, e)


def test_dense_scene_in_batch_equals_alone():
    dense = dense_scene()
    batch = [scenes.random_box_scene(i, 300, n_range=(500, 4000)) for i in range(3)]
    for sc in batch:
        sc["dt"] = dense["dt"]
    batch.insert(1, dense)
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    n = st["x"].shape[1]
    ctx = Context.scenes(descs, n)
    ctx.set_materials(mats)
    ctx.import_state({k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in st.items()})
    ctx.substeps(None, 10)
    allx = export(ctx)
    ctx.close()
    ctx = make_ctx(dense)
    ctx.substeps(params(dense), 10)
    alone = export(ctx)
    ctx.close()
    a, b = ranges[1]
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(allx[k][:, a:b], alone[k]), k


def test_parts_match_oracle(oracle):
    cfg = dense_scene()
    ctx = make_ctx(cfg)
    ctx.substeps(params(cfg), 30)
    a = export(ctx)
    ctx.close()
    assert_oracle(a, cfg, 30, oracle)


def test_parts_aflip_match_oracle(oracle):
    cfg = dense_scene()
    ctx = make_ctx(cfg, TRANSFER_AFLIP)
    ctx.substeps(params(cfg), 20)
    a = export(ctx)
    ctx.close()
    assert_oracle(a, cfg, 20, oracle, TRANSFER_AFLIP)


def test_fused_toggle_between_calls(oracle):
    """The fused chain needs one part per block: switching it on re-cuts the partition."""
    cfg = dense_scene()
    p = params(cfg)
    ctx = make_ctx(cfg)
    for k, fused in enumerate((False, True, False, True, False)):
        ctx.set_fused(fused)
        ctx.substeps(p, 6)
        if k == 1:  # a fused graph captured here runs after the unfused substeps of k = 2 .. 4
            ctx.graph_capture(p, 2)
    ctx.set_fused(False)
    ctx.graph_launch()
    a = export(ctx)
    ctx.close()
    assert_oracle(a, cfg, 32, oracle)
