"""GPU parity: libzmpm (HIP, gfx950) against the CPU oracle on the same seeded inputs.

Bars (stated per test):
  binning keys, sort order, partition order ..... bit-exact
  per-stage grid / particle values .............. FP32 tolerance (summation order, FMA)
  fused substeps vs the node-by-node chain ...... bit-exact (same kernels, same arithmetic)
  repeated runs ................................. bit-exact (no float atomics)
  N substeps vs FP32 oracle ..................... |dx| / dx <= 1e-3 cells (max), |dv| <= 1e-3 |v|inf
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from zeno_amd import scenes  # noqa: E402
from zeno_amd.zmpm import Context, StepParams, ZmpmError, particle_view  # noqa: E402
from zeno_amd.zmpm import LAYOUT_AOS, LAYOUT_AOSOA32  # noqa: E402

DEV = "cuda:0"


def to_dev(state):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(DEV) for k, v in state.items()}


def make_ctx(cfg, capacity=None):
    n = cfg["state"]["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], capacity or n)
    ctx.set_materials(cfg["mats"])
    ctx.import_state(to_dev(cfg["state"]))
    return ctx


def export_np(ctx):
    out = ctx.export_state(DEV)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def lattice_key(c, n_cells):
    nb = (n_cells + 3) // 4 + 3
    b = (c >> 2) + 1
    li = (b[0].astype(np.int64) * nb + b[1]) * nb + b[2]
    cell = ((c[0] & 3) * 4 + (c[1] & 3)) * 4 + (c[2] & 3)
    return ((li << 6) | cell).astype(np.uint32)


def grid_dict(keys, data):
    return {tuple(k): d for k, d in zip(keys.tolist(), data)}


# ------------------------------------------------------------------ binning (bit-exact)
def test_bin_keys_bit_exact(oracle):
    n_cells = 128
    dx = 1.0 / n_cells
    rng = np.random.default_rng(0)
    nb = (n_cells + 3) // 4 + 3
    hi_cell = 4 * (nb - 2)  # the lattice holds particle cells [-4, 4 (nb - 2)) on each axis
    x = rng.uniform(-3.5 * dx, (hi_cell - 0.5) * dx, size=(3, 200000)).astype(np.float32)
    dxinv = np.float32(1.0) / np.float32(dx)
    k = np.arange(-3, hi_cell - 1, dtype=np.float32)
    xb = ((k + np.float32(0.5)) / dxinv).astype(np.float32)
    x = np.concatenate([x, np.stack([xb] * 3), np.stack([np.nextafter(xb, -1)] * 3)], axis=1)
    ctx = Context(dx, n_cells, x.shape[1])
    keys = ctx.bin_keys(torch.from_numpy(x).to(DEV)).cpu().numpy().view(np.uint32)
    expect = lattice_key(oracle.cell_coords(x, dx), n_cells)
    assert np.array_equal(keys, expect)


def test_bin_keys_domain_error():
    ctx = Context(1 / 64, 64, 16)
    x = torch.full((3, 4), 0.5, device=DEV)
    x[0, 2] = 7.0
    with pytest.raises(ZmpmError) as e:
        ctx.bin_keys(x)
    assert e.value.code == 4
    x[0, 2] = float("nan")
    with pytest.raises(ZmpmError) as e:
        ctx.bin_keys(x)
    assert e.value.code == 4


@pytest.mark.parametrize("n", [1, 2, 100, 2047, 2048, 2049, 100003, 1 << 20])
def test_sort_keys_stable(n):
    rng = np.random.default_rng(n)
    ctx = Context(1 / 64, 64, max(n, 1))
    for bits, hi in ((24, 1 << 24), (25, 37), (9, 300)):
        keys = rng.integers(0, hi, size=n, dtype=np.int64).astype(np.uint32)
        ko, po = ctx.sort_keys(torch.from_numpy(keys.view(np.int32)).to(DEV), bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(po.cpu().numpy(), order)
        assert np.array_equal(ko.cpu().numpy().view(np.uint32), keys[order])


def test_partition_order_bit_exact(oracle):
    cfg = scenes.small_mixed_scene()
    st = cfg["state"]
    perm = np.random.default_rng(4).permutation(st["x"].shape[1])
    cfg["state"] = {k: np.ascontiguousarray(v[..., perm]) for k, v in st.items()}
    ctx = make_ctx(cfg)
    keys = lattice_key(oracle.cell_coords(cfg["state"]["x"], cfg["dx"]), cfg["n_cells"])
    # particles are stably grouped by block key (the cell bits do not take part)
    assert np.array_equal(ctx.order(), np.argsort(keys >> 6, kind="stable"))
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    o.stage("partition")
    assert ctx.num_grid_blocks() == oracle.lib().orc_num_gblocks_f32(o._h)
    assert ctx.num_particle_blocks() == oracle.lib().orc_num_pblocks_f32(o._h)


# ------------------------------------------------------------------ stages (FP32 tolerance)
def test_stage_parity(oracle):
    cfg = scenes.small_mixed_scene()
    dt = cfg["dt"]
    ctx = make_ctx(cfg)
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], dt)
    o.stage("partition")
    ctx.partition()
    ctx.grid_clear()
    ctx.p2g(dt)
    o.stage("p2g")
    gk, gd = ctx.grid()
    ok, od = o.grid()
    assert sorted(map(tuple, gk.tolist())) == sorted(map(tuple, ok.tolist()))
    G, O = grid_dict(gk, gd), grid_dict(ok, od)
    scale_m = max(abs(od[:, 0]).max(), 1e-30)
    scale_p = max(abs(od[:, 1:]).max(), 1e-30)
    for key, d in O.items():
        g = G[key]
        assert np.allclose(g[0], d[0], rtol=1e-5, atol=1e-6 * scale_m)
        assert np.allclose(g[1:], d[1:], rtol=1e-4, atol=2e-5 * scale_p)
    # grid update + boundary
    ctx.grid_update(dt, cfg["gravity"])
    o.stage("grid_update")
    assert np.isclose(ctx.max_vel_sqr(), o.max_vel_sqr(), rtol=1e-4)
    ctx.boundary(1, cfg["pad"])
    o.stage("boundary")
    gk, gd = ctx.grid()
    ok, od = o.grid()
    G, O = grid_dict(gk, gd), grid_dict(ok, od)
    vscale = max(abs(od[:, 1:]).max(), 1e-30)
    for key, d in O.items():
        assert np.allclose(G[key][1:], d[1:], rtol=1e-4, atol=1e-5 * vscale)
    # G2P
    ctx.g2p(dt)
    o.stage("g2p")
    a = export_np(ctx)
    b = o.state()
    assert np.abs(a["x"] - b["x"]).max() <= 1e-6
    vs = max(np.abs(b["v"]).max(), 1e-6)
    assert np.abs(a["v"] - b["v"]).max() <= 1e-4 * vs
    assert np.abs(a["F"] - b["F"]).max() <= 1e-5


# ------------------------------------------------------------------ chains
def run_gpu_steps(cfg, steps, fused=True):
    ctx = make_ctx(cfg)
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
    ctx.sync()
    return ctx, export_np(ctx)


def test_fused_equals_node_chain():
    cfg = scenes.small_mixed_scene()
    _, a = run_gpu_steps(cfg, 12, fused=True)
    _, b = run_gpu_steps(cfg, 12, fused=False)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_fast_particles_fused_equals_node_chain(oracle):
    """A solid box moving about half a cell per substep: many particles change block every
    substep, most are far from a block face.  The fused chain (which interpolates only the
    particles within reach of a face to find their new block) equals the node-by-node chain
    (full re-binning) bit for bit, and its resident order stays the stable block order."""
    cfg = scenes.small_mixed_scene(v_solid=(40.0, -60.0, 25.0))
    ctx, a = run_gpu_steps(cfg, 12, fused=True)
    _, b = run_gpu_steps(cfg, 12, fused=False)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    order = ctx.order()
    x_sorted = a["x"][:, order]
    keys = lattice_key(oracle.cell_coords(np.ascontiguousarray(x_sorted), cfg["dx"]), cfg["n_cells"]) >> 6
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


def test_incremental_binning_order(oracle):
    """After fused substeps (incremental binning) the resident order is the stable block order."""
    cfg = scenes.small_mixed_scene()
    ctx, a = run_gpu_steps(cfg, 10)
    order = ctx.order()
    x_sorted = a["x"][:, order]
    keys = lattice_key(oracle.cell_coords(np.ascontiguousarray(x_sorted), cfg["dx"]), cfg["n_cells"]) >> 6
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(order), np.arange(order.size))


def test_deterministic_repeat():
    cfg = scenes.small_mixed_scene()
    _, a = run_gpu_steps(cfg, 15)
    _, b = run_gpu_steps(cfg, 15)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_substeps_parity_mixed(oracle):
    cfg = scenes.small_mixed_scene()
    steps = cfg["steps"]
    _, a = run_gpu_steps(cfg, steps)
    o = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    o.substeps(steps)
    b = o.state()
    dx = cfg["dx"]
    assert np.abs(a["x"] - b["x"]).max() / dx <= 1e-3
    assert np.abs(a["v"] - b["v"]).max() <= 1e-3 * max(np.abs(b["v"]).max(), 1e-3)


def test_layout_round_trip():
    cfg = scenes.small_mixed_scene()
    st = cfg["state"]
    n = st["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    d = to_dev(st)
    J = d["F"][0].contiguous()
    # AoS import
    aos = {k: d[k].t().contiguous() for k in ("x", "v", "C", "F")}
    ctx.import_particles(particle_view(n, aos["x"], aos["v"], aos["C"], aos["F"], J, d["mat"], layout=LAYOUT_AOS))
    a = export_np(ctx)
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(a[k], st[k]), k
    # AoSoA32 import: the zeno TileVector channel map of an FCR "mpm" object (m0 x1 v4 vol7 C8 vms17 beta18 F19)
    nch = 28
    tiles = torch.zeros(((n + 31) // 32) * nch * 32, device=DEV)
    idx = torch.arange(n, device=DEV)
    chn = {"x": 1, "v": 4, "C": 8, "F": 19}
    for name, c0 in chn.items():
        src = d[name]
        for c in range(src.shape[0]):
            tiles[((idx // 32) * nch + c0 + c) * 32 + idx % 32] = src[c]
    Jt = torch.zeros(((n + 31) // 32) * nch * 32, device=DEV)
    Jt[((idx // 32) * nch + 19) * 32 + idx % 32] = J
    ctx.import_particles(particle_view(n, tiles, tiles, tiles, tiles, Jt, d["mat"], layout=LAYOUT_AOSOA32,
                                       tile_channels=nch, chn=dict(chn, J=19)))
    b = export_np(ctx)
    for k in ("x", "v", "C", "F"):
        assert np.array_equal(b[k], st[k]), k


def test_plain_cube_parity(oracle):
    """65,536-particle cube (dx = 1/128, FCR E = 5e4, nu = 0.2), 200 substeps: FP32 GPU vs FP32 and FP64 oracle."""
    cfg = scenes.config_plain_cube()
    steps = cfg["steps"]
    _, a = run_gpu_steps(cfg, steps)
    o32 = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"])
    o32.substeps(steps)
    b = o32.state()
    o64 = oracle.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"], dtype=np.float64)
    o64.substeps(steps)
    c = o64.state()
    dx = cfg["dx"]
    ex32 = np.linalg.norm(a["x"] - b["x"], axis=0) / dx
    ex64 = np.linalg.norm(a["x"] - c["x"], axis=0) / dx
    vinf = np.abs(c["v"]).max()
    ev32 = np.linalg.norm(a["v"] - b["v"], axis=0) / vinf
    print(f"\nplain cube 200 steps: |dx|/dx vs f32 max {ex32.max():.3e} rms {np.sqrt((ex32**2).mean()):.3e}; "
          f"vs f64 max {ex64.max():.3e}; |dv|/|v|inf vs f32 max {ev32.max():.3e}")
    assert ex32.max() <= 1e-3
    assert ev32.max() <= 1e-3
    assert ex64.max() <= 1e-3


# ------------------------------------------------------------------ particle volume (bit-exact)
def test_particle_volume_bit_exact(oracle):
    """ComputeParticleVolume: vol = dx^3 / (particles in the cell floor(x/dx)), m = rho vol, per
    caller particle index, bit-exact against the oracle (integer counts, one correctly rounded
    division, one product), before and after substeps reorder the resident state."""
    rng = np.random.default_rng(21)
    n_cells = 64
    dx = 1.0 / n_cells
    n = 60000
    centers = rng.uniform(0.35, 0.65, size=(3, 12))
    pick = rng.integers(0, 12, n)
    x = (centers[:, pick] + 0.02 * rng.standard_normal((3, n))).astype(np.float32)
    mat = rng.integers(0, 2, n).astype(np.int32)
    mats = [dict(model=scenes.MODEL_FCR, E=1e4, nu=0.3, rho=800.0, vol=dx ** 3 / 8),
            dict(model=scenes.MODEL_FCR, E=2e4, nu=0.2, rho=1700.0, vol=dx ** 3 / 8)]
    st = scenes._state(x, (0.0, 0.0, 0.0), mat, np.array([scenes.MODEL_FCR, scenes.MODEL_FCR]))
    cfg = dict(state=st, mats=mats, dx=dx, n_cells=n_cells)
    ctx = make_ctx(cfg)
    rho = np.array([800.0, 1700.0], np.float32)[mat]
    for step in range(2):
        xs = export_np(ctx)["x"]
        vol, m = ctx.particle_volume(DEV)
        ev, em = oracle.particle_volume(xs, dx, rho)
        assert np.array_equal(vol.cpu().numpy(), ev), step
        assert np.array_equal(m.cpu().numpy(), em), step
        ctx.substeps(StepParams(1e-4, (0.0, -9.8, 0.0), 1, 3), 5)
    ctx.close()


def test_particle_volume_into_tilevector():
    """Written into a zpc TileVector<f32,32> (AoSoA32, the "mpm" map: m = channel 0, vol = 7)
    the values equal the SoA output, and no other channel is touched."""
    rng = np.random.default_rng(5)
    n = 4099
    x = (0.5 + 0.05 * rng.standard_normal((3, n))).astype(np.float32)
    mats = [dict(model=scenes.MODEL_FCR, E=1e4, nu=0.3, rho=1200.0, vol=1e-6)]
    st = scenes._state(x, (0.0, 0.0, 0.0), np.zeros(n, np.int32), np.array([scenes.MODEL_FCR]))
    ctx = make_ctx(dict(state=st, mats=mats, dx=1.0 / 128, n_cells=128))
    vol, m = ctx.particle_volume(DEV)
    tiles = (n + 31) // 32
    tv = torch.full((tiles * 28 * 32,), -7.0, device=DEV)
    ctx.particle_volume_into(LAYOUT_AOSOA32, 28, tv, 7, tv, 0)
    ctx.sync()
    t = tv.view(tiles, 28, 32).permute(1, 0, 2).reshape(28, -1)[:, :n]
    assert torch.equal(t[7], vol) and torch.equal(t[0], m)
    others = [c for c in range(28) if c not in (0, 7)]
    assert bool((t[others] == -7.0).all())
    with pytest.raises(ZmpmError, match="ZMPM_ERR_ARG"):
        ctx.particle_volume_into(LAYOUT_AOSOA32, 28, tv, 28, None, 0)
    ctx.close()


def test_particle_volume_scenes_bit_exact(oracle):
    """In a multi-scene context each scene counts its own cells with its own dx."""
    batch = [scenes.random_box_scene(i, 4, n_range=(2048, 8192)) for i in range(5)]
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    n = st["x"].shape[1]
    ctx = Context.scenes(descs, n)
    ctx.set_materials(mats)
    ctx.import_state(to_dev(st))
    vol, m = ctx.particle_volume(DEV)
    group = np.concatenate([np.full(b - a, k) for k, (a, b) in enumerate(ranges)])
    dxp = np.concatenate([np.full(b - a, batch[k]["dx"], np.float32) for k, (a, b) in enumerate(ranges)])
    rho = np.array([mt["rho"] for mt in mats], np.float32)[st["mat"]]
    ev, em = oracle.particle_volume(st["x"], dxp, rho, group=group)
    assert np.array_equal(vol.cpu().numpy(), ev)
    assert np.array_equal(m.cpu().numpy(), em)
    ctx.close()
