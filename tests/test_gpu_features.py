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
