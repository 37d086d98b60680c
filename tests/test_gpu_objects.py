"""The "ZSParticles" input as the reference hands it over, and the per-particle mass and volume.

  objects   a J-fluid TileVector (20 channels, J at 19) and a fixed-corotated one (28 channels,
            F at 19), the "mpm" channel maps of Generation.cpp:868-884, imported in place as two
            objects (zmpm_import_objects): export round trip bit-exact (channels not handed over
            keep the caller's values), 20 substeps vs the FP32 oracle at 1e-3 cells.
  volume    ComputeParticleVolume's m, vol reach P2G (Maintenance.cu:50-51 -> Transfer.cu:857-858):
            after zmpm_particle_volume the substeps equal the oracle run with those per-particle
            m, vol (1e-3 cells) and differ from the run without them; imported m, vol likewise.
  CFL       every fused substep (eager and graph replays) leaves UpdateZSGrid's MaxVelSqr on the
            device (zmpm_max_vel_sqr_device), equal to the synchronising zmpm_max_vel_sqr.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from zeno_amd import scenes
from zeno_amd.zmpm import (LAYOUT_AOSOA32, MODEL_FCR, MODEL_JFLUID, Context, StepParams, particle_view)

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(st):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in st.items()}


def tilevector(fields, nch, n, sentinel=-7.0):
    """AoSoA32 TileVector<f32, 32> of n particles, nch channels: element (chn, i) at
    ((i / 32) nch + chn) 32 + i % 32.  fields: {name: (chn, array (k, n))}."""
    ntiles = (n + 31) // 32
    tv = np.full((ntiles, nch, 32), sentinel, np.float32)
    i = np.arange(n)
    for chn, arr in fields.values():
        for c in range(arr.shape[0]):
            tv[i // 32, chn + c, i % 32] = arr[c]
    return tv


def tv_channel(tv, chn, width, n):
    i = np.arange(n)
    return np.stack([tv[i // 32, chn + c, i % 32] for c in range(width)])


FCR_MAP = {"m": 0, "x": 1, "v": 4, "vol": 7, "C": 8, "vms": 17, "beta": 18, "F": 19}   # 28 channels
FLUID_MAP = {"m": 0, "x": 1, "v": 4, "vol": 7, "C": 8, "vms": 17, "beta": 18, "J": 19}  # 20 channels


def _objects():
    cfg = scenes.small_mixed_scene()
    st = cfg["state"]
    solid = st["mat"] == 0
    fluid = st["mat"] == 1
    rng = np.random.default_rng(5)
    n = st["x"].shape[1]
    C = (0.5 * rng.standard_normal((9, n))).astype(np.float32)
    Fs = (np.eye(3, dtype=np.float32).reshape(9, 1) + 0.02 * rng.standard_normal((9, n))).astype(np.float32)
    J = (1.0 + 0.01 * rng.standard_normal(n)).astype(np.float32)
    ns, nf = int(solid.sum()), int(fluid.sum())
    tv_s = tilevector({"x": (1, st["x"][:, solid]), "v": (4, st["v"][:, solid]), "C": (8, C[:, solid]),
                       "F": (19, Fs[:, solid])}, 28, ns)
    tv_f = tilevector({"x": (1, st["x"][:, fluid]), "v": (4, st["v"][:, fluid]), "C": (8, C[:, fluid]),
                       "J": (19, J[None, fluid])}, 20, nf)
    # the oracle's state: object 0 (solid) then object 1 (fluid), in the objects' order
    order = np.concatenate([np.nonzero(solid)[0], np.nonzero(fluid)[0]])
    F = np.zeros((9, n), np.float32)
    F[:, solid] = Fs[:, solid]
    F[0, fluid] = J[fluid]
    ref = {"x": st["x"][:, order], "v": st["v"][:, order], "C": C[:, order], "F": F[:, order],
           "mat": st["mat"][order]}
    return cfg, tv_s, tv_f, ns, nf, ref


def _views(t_s, t_f, ns, nf):
    vs = particle_view(ns, t_s, v=t_s, C=t_s, F=t_s, layout=LAYOUT_AOSOA32, tile_channels=28,
                       chn=FCR_MAP, default_material=0)
    vf = particle_view(nf, t_f, v=t_f, C=t_f, J=t_f, layout=LAYOUT_AOSOA32, tile_channels=20,
                       chn=FLUID_MAP, default_material=1)
    return [vs, vf]


def test_import_objects_tilevectors_round_trip_and_parity():
    cfg, tv_s, tv_f, ns, nf, ref = _objects()
    assert cfg["mats"][0]["model"] == MODEL_FCR and cfg["mats"][1]["model"] == MODEL_JFLUID
    n = ns + nf
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    t_s, t_f = torch.from_numpy(tv_s).to(DEV), torch.from_numpy(tv_f).to(DEV)
    ctx.import_objects(_views(t_s, t_f, ns, nf))
    # export into fresh TileVectors: the handed-over channels come back bit for bit, the rest
    # (m, vol, vms, beta) keep the caller's values
    o_s = torch.full_like(t_s, 3.5)
    o_f = torch.full_like(t_f, 3.5)
    ctx.export_objects(_views(o_s, o_f, ns, nf))
    h_s, h_f = o_s.cpu().numpy(), o_f.cpu().numpy()
    for chn, w in ((1, 3), (4, 3), (8, 9), (19, 9)):
        np.testing.assert_array_equal(tv_channel(h_s, chn, w, ns), tv_channel(tv_s, chn, w, ns))
    for chn, w in ((1, 3), (4, 3), (8, 9), (19, 1)):
        np.testing.assert_array_equal(tv_channel(h_f, chn, w, nf), tv_channel(tv_f, chn, w, nf))
    for chn in (0, 7, 17, 18):
        assert np.all(tv_channel(h_s, chn, 1, ns) == 3.5) and np.all(tv_channel(h_f, chn, 1, nf) == 3.5)
    # 20 substeps against the oracle on the same particles
    steps = 20
    p = StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])
    ctx.substeps(p, steps)
    ctx.export_objects(_views(o_s, o_f, ns, nf))
    ctx.sync()
    h_s, h_f = o_s.cpu().numpy(), o_f.cpu().numpy()
    x = np.concatenate([tv_channel(h_s, 1, 3, ns), tv_channel(h_f, 1, 3, nf)], axis=1)
    v = np.concatenate([tv_channel(h_s, 4, 3, ns), tv_channel(h_f, 4, 3, nf)], axis=1)
    J = tv_channel(h_f, 19, 1, nf)[0]
    ctx.close()
    s = O.Session(ref, cfg["mats"], cfg["dx"], cfg["dt"], gravity=cfg["gravity"], pad=cfg["pad"])
    s.substeps(steps)
    r = s.state()
    assert np.abs(x - r["x"]).max() / cfg["dx"] < 1e-3
    assert np.abs(v - r["v"]).max() < 1e-3 * np.abs(r["v"]).max()
    assert np.abs(J - r["F"][0, ns:]).max() < 1e-4


def _uneven_scene():
    """A solid block at 8 particles per cell next to one at 2, and a fluid slab whose cells hold
    1 to 6 particles: per-cell volumes differ from the materials' dx^3 / ppc."""
    dx = 1.0 / 32
    a = scenes.jitter_cells(scenes.box_cells((6, 10, 6), (10, 14, 10)), 8, dx, 31)
    b = scenes.jitter_cells(scenes.box_cells((10, 10, 6), (14, 14, 10)), 2, dx, 32)
    cells = scenes.box_cells((16, 4, 8), (24, 8, 16))
    rng = np.random.default_rng(33)
    counts = rng.integers(1, 7, len(cells))
    c = np.concatenate([scenes.jitter_cells(cells[k:k + 1], int(counts[k]), dx, 34, counter0=100 * k)
                        for k in range(len(cells))], axis=1)
    mats = [dict(model=MODEL_FCR, E=3e4, nu=0.3, rho=1000.0, vol=dx ** 3 / 4),
            dict(model=MODEL_JFLUID, rho=800.0, vol=dx ** 3 / 4, bulk=4e3)]
    model_of = np.array([MODEL_FCR, MODEL_JFLUID])
    st = scenes.concat_states([scenes._state(np.concatenate([a, b], 1), (0.2, -0.4, 0.1),
                                             np.zeros(a.shape[1] + b.shape[1], np.int32), model_of),
                               scenes._state(c, (0.0, 0.0, 0.0), np.ones(c.shape[1], np.int32), model_of)])
    # pre-strained, so that the volume-weighted stress acts from the first substep
    st["F"] = (st["F"] * np.float32(1.03)).astype(np.float32)
    st["F"][0] = np.where(st["mat"] == 1, np.float32(0.97), st["F"][0])
    return st, mats, dx


def test_particle_volume_reaches_p2g():
    st, mats, dx = _uneven_scene()
    n = st["x"].shape[1]
    dt, steps = 2e-4, 20
    p = StepParams(dt, (0.0, -9.8, 0.0), 1, 3)
    ctx = Context(dx, 32, n)
    ctx.set_materials(mats)
    ctx.import_state(_dev(st))
    vol, mass = ctx.particle_volume(DEV)
    vol, mass = vol.cpu().numpy(), mass.cpu().numpy()
    rho = np.array([m["rho"] for m in mats], np.float32)[st["mat"]]
    ov, om = O.particle_volume(st["x"], dx, rho, group=None)
    np.testing.assert_array_equal(vol, ov)
    np.testing.assert_array_equal(mass, om)
    ctx.substeps(p, steps)
    out = {k: t.cpu().numpy() for k, t in ctx.export_state(DEV, aux=True).items()}
    ctx.close()
    # the particles keep their own m, vol, exported as such
    np.testing.assert_array_equal(out["vol"], ov)
    np.testing.assert_array_equal(out["m"], om)
    s = O.Session(st, mats, dx, dt, gravity=(0.0, -9.8, 0.0), pad=3)
    s.set_mass_vol(om, ov)
    s.substeps(steps)
    r = s.state()
    assert np.abs(out["x"] - r["x"]).max() / dx < 1e-3
    assert np.abs(out["v"] - r["v"]).max() < 1e-3 * np.abs(r["v"]).max()
    # without the per-particle values the run is a different one
    s0 = O.Session(st, mats, dx, dt, gravity=(0.0, -9.8, 0.0), pad=3)
    s0.substeps(steps)
    assert np.abs(s0.state()["x"] - r["x"]).max() / dx > 1e-2
    # imported per-particle m, vol give the same run bit for bit
    st2 = dict(st, m=om, vol=ov)
    c2 = Context(dx, 32, n)
    c2.set_materials(mats)
    c2.import_state(_dev(st2))
    c2.substeps(p, steps)
    out2 = {k: t.cpu().numpy() for k, t in c2.export_state(DEV).items()}
    c2.close()
    for k in ("x", "v", "C", "F"):
        np.testing.assert_array_equal(out2[k], out[k])


def _read_device_floats(addr, count):
    hip = ctypes.CDLL("libamdhip64.so")
    out = (ctypes.c_float * count)()
    rc = hip.hipMemcpy(out, ctypes.c_void_p(addr), ctypes.c_size_t(4 * count), ctypes.c_int(2))  # device to host
    assert rc == 0
    return np.array(out[:], np.float32)


def test_max_vel_sqr_every_substep():
    cfg = scenes.small_mixed_scene()
    st = cfg["state"]
    n = st["x"].shape[1]
    p = StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    ctx.import_state(_dev(st))
    addr = ctx.max_vel_sqr_device()
    assert addr != 0
    s = O.Session(st, cfg["mats"], cfg["dx"], cfg["dt"])
    for k in range(6):
        ctx.substeps(p, 1)
        ctx.sync()
        dev = _read_device_floats(addr, 1)[0]
        assert dev == np.float32(ctx.max_vel_sqr()), k
        s.substeps(1)
        assert abs(dev - s.max_vel_sqr()) <= 1e-4 * s.max_vel_sqr(), k
    ctx.close()
    # graph replays of a multi-scene context write every scene's maximum
    batch = [scenes.random_box_scene(i, 300, n_range=(1000, 3000)) for i in range(3)]
    descs, mats, bst, _ = scenes.merge_scenes(batch)
    g = Context.scenes(descs, bst["x"].shape[1])
    g.set_materials(mats)
    g.import_state(_dev(bst))
    g.graph_capture(None, 2)
    addr = g.max_vel_sqr_device()
    for rep in range(3):
        g.graph_launch()
        g.sync()
        torch.cuda.synchronize()
        dev = _read_device_floats(addr, 3)
        np.testing.assert_array_equal(dev, g.scene_max_vel_sqr())
        assert np.all(dev > 0)
    g.close()
