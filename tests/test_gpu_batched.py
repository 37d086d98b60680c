"""Many independent scenes in one context (every substep launch covers all of them).

Each scene of a batch must evolve exactly as it does alone: the kernels treat a
block the same whatever else shares the launch, the binning keeps each scene's
particles in their own part of the lattice, and every scene uses its own dx,
dt, gravity and walls.  So a batched run is compared bit for bit with one
single-scene context per scene (those are compared with the oracle in
test_gpu_parity.py).  HIP graph replays are compared with eager substeps,
with eager work and synchronisation between the replays.
"""
import numpy as np
import pytest
import torch

from zeno_amd import scenes
from zeno_amd.zmpm import BOUNDARY_STICKY_BOX, Context, StepParams, ZmpmError

pytestmark = pytest.mark.gpu


def _dev(st):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for k, v in st.items()}


def _batch():
    out = [scenes.random_box_scene(i, 100, n_range=(300, 3000)) for i in range(4)]
    mixed = scenes.small_mixed_scene()  # fixed corotated + J-fluid, n_cells 32, dt 2e-4
    out.insert(2, mixed)
    return out


def _single(sc, nsteps):
    n = sc["state"]["x"].shape[1]
    ctx = Context(sc["dx"], sc["n_cells"], n)
    ctx.set_materials(sc["mats"])
    ctx.import_state(_dev(sc["state"]))
    ctx.substeps(StepParams(sc["dt"], sc["gravity"], BOUNDARY_STICKY_BOX, sc["pad"]), nsteps)
    out = ctx.export_state()
    vm = ctx.max_vel_sqr()
    ctx.close()
    return out, vm


def _batched_ctx(batch):
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    n = st["x"].shape[1]
    ctx = Context.scenes(descs, n)
    ctx.set_materials(mats)
    ctx.import_state(_dev(st))
    return ctx, ranges


def test_batched_equals_one_context_per_scene():
    batch = _batch()
    ctx, ranges = _batched_ctx(batch)
    assert ctx.num_scenes() == len(batch)
    ctx.substeps(None, 12)
    out = ctx.export_state()
    vms = ctx.scene_max_vel_sqr()
    assert ctx.max_vel_sqr() == vms.max()
    for i, sc in enumerate(batch):
        ref, vm = _single(sc, 12)
        a, b = ranges[i]
        for k in ("x", "v", "C", "F"):
            assert torch.equal(out[k][:, a:b], ref[k]), (i, k)
        assert vms[i] == np.float32(vm), i
    ctx.close()


def test_graph_replays_match_eager_substeps():
    batch = _batch()
    eager, _ = _batched_ctx(batch)
    graph, _ = _batched_ctx(batch)
    graph.graph_capture(None, 6)
    for rep in range(3):
        eager.substeps(None, 6)
        graph.graph_launch()
        e = eager.export_state()
        g = graph.export_state()
        for k in ("x", "v", "C", "F"):
            assert torch.equal(e[k], g[k]), (rep, k)
        assert np.array_equal(eager.scene_max_vel_sqr(), graph.scene_max_vel_sqr()), rep
    eager.close()
    graph.close()


def test_graph_with_step_params_single_scene():
    sc = scenes.small_mixed_scene()
    n = sc["state"]["x"].shape[1]
    p = StepParams(sc["dt"], sc["gravity"], BOUNDARY_STICKY_BOX, sc["pad"])
    ctxs = []
    for _ in range(2):
        c = Context(sc["dx"], sc["n_cells"], n)
        c.set_materials(sc["mats"])
        c.import_state(_dev(sc["state"]))
        ctxs.append(c)
    ctxs[1].graph_capture(p, 4)
    for rep in range(2):
        ctxs[0].substeps(p, 4)
        ctxs[1].graph_launch()
        a, b = ctxs[0].export_state(), ctxs[1].export_state()
        for k in ("x", "v", "C", "F"):
            assert torch.equal(a[k], b[k]), (rep, k)
    for c in ctxs:
        c.close()


def test_scene_errors():
    batch = _batch()[:2]
    descs, mats, st, _ = scenes.merge_scenes(batch)
    ctx = Context.scenes(descs, st["x"].shape[1])
    bad = [dict(m, scene=5) for m in mats]
    with pytest.raises(ZmpmError, match="ZMPM_ERR_ARG"):
        ctx.set_materials(bad)
    ctx.set_materials(mats)
    ctx.import_state(_dev(st))
    with pytest.raises(ZmpmError, match="ZMPM_ERR_ARG"):
        ctx.graph_capture(None, 3)
    with pytest.raises(ZmpmError, match="ZMPM_ERR_STATE"):
        ctx.graph_launch()
    ctx.close()
    with pytest.raises(ZmpmError, match="ZMPM_ERR_ARG"):
        Context.scenes([], 10)


def test_autotune_keeps_state_and_results():
    """zmpm_autotune times the workgroup-size candidates (and the fused pass, update size 0) from
    the current state, restores it, and the chosen sizes give substeps bit-identical to an untuned
    context with the same fused setting (which fixes how P2G groups the sums of blocks of more than
    one chunk)."""
    batch = _batch()
    plain, _ = _batched_ctx(batch)
    tuned, _ = _batched_ctx(batch)
    choice = tuned.autotune(None, 2)
    assert choice[0] in (64, 256) and choice[1] in (0, 64, 256) and choice[2] in (True, False)
    assert tuned.fused() == (choice[1] == 0)
    plain.set_fused(tuned.fused())
    for k in ("x", "v", "C", "F"):
        assert torch.equal(plain.export_state()[k], tuned.export_state()[k]), k
    plain.substeps(None, 6)
    tuned.substeps(None, 6)
    a, b = plain.export_state(), tuned.export_state()
    for k in ("x", "v", "C", "F"):
        assert torch.equal(a[k], b[k]), k
    plain.close()
    tuned.close()
