#!/usr/bin/env python3
"""Throughput benchmark of the MLS/APIC-MPM substep on one MI355X.

Workload (the throughput configuration): 4,194,304 particles at dx = 1/256 --
a J-fluid dam-break column (bulk modulus 1e4) plus a fixed-corotated elastic
sphere (r = 0.08, E = 1e5, v0 = (0, -2, 0)), jittered at 8 particles per cell,
sparse 4^3 blocks, dt = 5e-5.  One "step" is one full substep (binning, P2G,
grid update + sticky box + CFL max, G2P) over the resident particles.

Prints ONE JSON line: whole-job particle-substeps/s over exactly --steps timed
substeps after --warmup untimed ones, the roofline object of the dominant
kernel (HIP events on the library's stream), and the CPU baseline (the FP32 CPU
oracle on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s @4M particles (+% HBM roofline); batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# timed phase -> the kernel it launches (as rocprofv3 names it); G2P sizes come from the autotuner,
# whose update size 0 means the fused G2P / P2G pass (the stress and P2G phases then only hold
# the first substep's P2G of a call whose tiles were not carried over)
def kernel_names(wg_move=256, wg_update=256):
    if wg_update == 0:
        return {"stress": "k_stress", "p2g": "k_p2g<0, false, true>", "grid": "k_grid",
                "g2p_move": f"k_g2p_move<{wg_move}, false, true>", "partition": "k_lattice_scan",
                "g2p_update": "k_g2p2g<false, false>"}
    return {"stress": "k_stress", "p2g": "k_p2g<0, false, false>", "grid": "k_grid",
            "g2p_move": f"k_g2p_move<{wg_move}>",
            "partition": "k_lattice_scan", "g2p_update": f"k_g2p_update<true, {wg_update}>"}


def algorithmic_bytes(n_solid, n_fluid, n_cells):
    """Per-phase algorithmic HBM bytes of one substep (FP32 state, int32 material id).

    Particle state: x3 v3 C9 + F9 (fixed corotated) or J1 (J-fluid), + material id.
      stress reads F (+ id), writes the 9-float stress term -> 72 B solid (+ 4 B id per particle)
      P2G reads x, v, C, stress|J, id           -> 100 B solid, 68 B fluid; writes grid m, mv: 16 B / active node
      grid update reads m, mv, writes v         -> 28 B / active node
      G2P reads x, F|J, id (+ grid v 12 B/node) -> 52 B solid, 20 B fluid; writes x, v, C, F|J: 96 B / 64 B
      G2P move reads x (+ grid v)               -> 12 B / particle + 12 B / active node
    Whole substep (fused P2G/G2P ideal): 196 B solid + 132 B fluid + 32 B per active node.
    Fused G2P / P2G pass: the particle state read once and written once (196 B solid, 132 B fluid),
      the updated grid velocities read (12 B / node) and the next P2G's grid written (16 B / node).
    """
    return {
        "g2p2g": 196 * n_solid + 132 * n_fluid + 28 * n_cells,
        "stress": (36 + 36) * n_solid + 4 * (n_solid + n_fluid),
        "p2g": 100 * n_solid + 68 * n_fluid + 16 * n_cells,
        "grid": 28 * n_cells,
        "g2p_update": (52 + 96) * n_solid + (20 + 64) * n_fluid + 12 * n_cells,
        "g2p_move": 12 * (n_solid + n_fluid) + 12 * n_cells,
        "substep": 196 * n_solid + 132 * n_fluid + 32 * n_cells,
    }


def pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, or None.

    The summary holds one JSON object between ```json fences with FETCH_SIZE and
    WRITE_SIZE per dispatch (KiB, as the counters read) and the copy-kernel
    calibration factors that turn them into bytes.
    """
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        blob = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        fetch = blob["fetch_kib_per_dispatch"] * 1024 * blob["fetch_calibration"]
        write = blob["write_kib_per_dispatch"] * 1024 * blob["write_calibration"]
    except (IndexError, KeyError, ValueError):
        return None
    return {"bytes": fetch + write, "fetch": fetch, "write": write, "source": os.path.relpath(path, ROOT)}


def copy_bandwidth(dev, nbytes=1 << 29, reps=5):
    """Device-to-device copy rate on this GPU: (read + write bytes) / time of a torch copy_ of
    nbytes, timed with events on torch's stream after one warm-up copy."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    return gbs


def cpu_baseline(cfg, steps):
    """FP32 CPU oracle (the C++/OpenMP restatement of the reference arithmetic) on this host."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    s = O.Session(cfg["state"], cfg["mats"], cfg["dx"], cfg["dt"], gravity=cfg["gravity"], pad=cfg["pad"],
                  threads=threads)
    n = cfg["state"]["x"].shape[1]
    t0 = time.perf_counter()
    s.substeps(steps)
    t = time.perf_counter() - t0
    s.close()
    return {"value": n * steps / t, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{cfg['name']}: {n} particles x {steps} substeps from the initial state, "
                      f"FP32 oracle (FP64 polar), OpenMP {threads} threads, {t:.2f} s"}


def batched_rates(n_scenes, launches, dev, one_by_one=True):
    """Batched workload: n_scenes random boxes (2k-64k particles, dx 1/64 or 1/128).

    Batched: one multi-scene context, 10 substeps captured in a HIP graph, `launches`
    timed graph launches after one untimed.  One by one: the same scenes in one
    single-scene context each (autotuned on its own), run one after another, 10 *
    launches substeps each (after 10 untimed), timed twice: as plain substep launches
    and as `launches` launches of the scene's own 10-substep graph; each set of times
    is summed.
    """
    import numpy as np
    import torch
    from zeno_amd import scenes
    from zeno_amd.zmpm import BOUNDARY_STICKY_BOX, Context, StepParams

    def to_dev(st):
        return {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in st.items()}

    batch = scenes.config_batched(n_scenes)
    descs, mats, st, _ = scenes.merge_scenes(batch)
    n_total = st["x"].shape[1]
    ctx = Context.scenes(descs, n_total)
    ctx.set_materials(mats)
    ctx.import_state(to_dev(st))
    wg = ctx.autotune(None, 2)
    ctx.graph_capture(None, 10)
    ctx.graph_launch()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        ctx.graph_launch()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    ctx.sync()
    ctx.close()
    del ctx
    t_single = t_graph = 0.0
    for sc in (batch if one_by_one else []):  # one scene per context and launch chain
        n = sc["state"]["x"].shape[1]
        c = Context(sc["dx"], sc["n_cells"], n)
        c.set_materials(sc["mats"])
        c.import_state(to_dev(sc["state"]))
        p = StepParams(sc["dt"], sc["gravity"], BOUNDARY_STICKY_BOX, sc["pad"])
        c.autotune(p, 2)
        c.substeps(p, 10)
        c.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c.substeps(p, 10 * launches)
        torch.cuda.synchronize()
        t_single += time.perf_counter() - t0
        c.graph_capture(p, 10)
        c.graph_launch()
        c.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(launches):
            c.graph_launch()
        torch.cuda.synchronize()
        t_graph += time.perf_counter() - t0
        c.sync()
        c.close()
    work = n_total * 10 * launches
    steps = 10 * launches
    return {"scenes": n_scenes, "particles": n_total, "substeps": steps,
            "batched": work / t_batched, "one_by_one": work / t_single if one_by_one else None,
            "ratio": t_single / t_batched if one_by_one else None,
            "one_by_one_graph": work / t_graph if one_by_one else None,
            "ratio_graph": t_graph / t_batched if one_by_one else None,
            "unit": "particle-substeps/s", "batched_ms_per_substep": t_batched / steps * 1e3,
            "one_by_one_ms_per_substep": t_single / steps * 1e3 if one_by_one else None,
            "one_by_one_graph_ms_per_substep": t_graph / steps * 1e3 if one_by_one else None,
            "autotuned": {"g2p_move_wg": wg[0], "g2p_update_wg": wg[1], "sparse_p2g": wg[2]},
            "method": "batched: one multi-scene context, 10-substep HIP graph; one by one: one context per "
                      "scene, plain substep launches and the scene's own 10-substep graph, times summed"}


def adjoint_rates(dev, checkpoint_every=8):
    """Differentiable workload: 131,072-particle elastic block (E 5e4, nu 0.3, dx 1/128),
    64 substeps at dt 1e-4.  Loss |x_cm(T) - x*|^2 (x* seeded in [0.3, 0.7]^3) plus the mean
    squared distance to final positions simulated with 1.5 E and nu + 0.05 (the centre of mass
    alone does not depend on E or nu).  Times the checkpointed forward and the reverse pass
    (recompute + adjoint substeps)."""
    import numpy as np
    import torch
    from zeno_amd import scenes
    from zeno_amd.diff import Rollout
    from zeno_amd.zmpm import BOUNDARY_STICKY_BOX, Context, StepParams

    cfg = scenes.config_diff_block()
    st = cfg["state"]
    n = st["x"].shape[1]
    target = 0.3 + 0.4 * scenes.uniform24(5, np.arange(3, dtype=np.uint64))
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    p = StepParams(cfg["dt"], cfg["gravity"], BOUNDARY_STICKY_BOX, cfg["pad"])
    sim = Rollout(ctx, st, cfg["mats"], p, cfg["steps"], checkpoint_every=checkpoint_every)
    v0 = torch.from_numpy(np.ascontiguousarray(st["v"])).to(dev)
    E, nu = cfg["mats"][0]["E"], cfg["mats"][0]["nu"]
    tgt = torch.tensor(target, dtype=torch.float32, device=dev)

    y = sim.run(v0, 1.5 * E, nu + 0.05).clone()

    def once():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = sim.run(v0, E, nu)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        d = x.mean(dim=1) - tgt
        xbar = (2.0 / n) * (d[:, None] + (x - y))
        g = sim.grad(xbar)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, g

    once()  # warm-up
    tf, tb, g = once()
    ctx.close()
    return {"workload": cfg["name"], "particles": n, "substeps": cfg["steps"], "checkpoint_every": checkpoint_every,
            "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "ratio": tb / tf,
            "grad_E": float(g["E"][0]), "grad_nu": float(g["nu"][0]),
            "method": "forward: substeps with a state checkpoint every k; backward: per segment one recompute "
                      "saving each substep's input state, then the adjoint substeps in reverse"}


def sysid_rates():
    """Identification workload: 64 scenes x 32,768 particles (dx 1/96), 48 substeps at dt 1e-4,
    50 batched Adam iterations on (ln E, nu) per scene from the adjoint gradients, all scenes in
    one multi-scene context.  Times the iterations (each: checkpointed forward + reverse pass)."""
    import numpy as np
    from zeno_amd import sysid
    batch = sysid.make_scenes()
    r = sysid.identify(batch, steps=48, iters=50)
    n, it, steps = r["particles"], r["iters"], r["steps"]
    return {"scenes": r["scenes"], "particles": n, "substeps": steps, "iterations": it,
            "seconds": r["seconds"], "ms_per_iteration": r["seconds"] / it * 1e3,
            "forward_particle_substeps_per_s": n * steps * it / r["seconds"],
            "loss_first": r["loss"][0], "loss_last": r["loss"][-1],
            "median_E_rel_err": float(np.median(r["E_rel_err"])), "median_nu_rel_err": float(np.median(r["nu_rel_err"])),
            "method": "each iteration: one checkpointed forward over all scenes, one reverse pass, one Adam "
                      "update of every scene's (ln E, nu); the rate counts forward particle-substeps only"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed substeps before timing")
    ap.add_argument("--cpu-steps", type=int, default=100, help="substeps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--batched-launches", type=int, default=3,
                    help="timed 10-substep graph launches of the 256-scene batched workload (0 = skip)")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched workload's timed graph launches (for profiling) and print its object")
    ap.add_argument("--one-by-one", type=int, default=1,
                    help="also time the batched scenes one by one (0 = skip; the ratio is then omitted)")
    ap.add_argument("--adjoint", type=int, default=1, help="time the differentiable workload (0 = skip)")
    ap.add_argument("--sysid", type=int, default=1, help="time the identification workload (0 = skip)")
    ap.add_argument("--fused", choices=("auto", "1", "0"), default="auto",
                    help="the fused G2P / P2G pass: the autotuner's choice, or forced on / off")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run a plain device-to-device copy of BYTES bytes (3 launches), the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the multi-GPU path (slab decomposition with halo exchange) "
              "is a design only in this build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from zeno_amd import scenes
    from zeno_amd.zmpm import Context, StepParams

    dev = "cuda:0"
    torch.cuda.set_device(0)
    if args.batched_only:
        print(json.dumps({"batched": batched_rates(args.batched_scenes, args.batched_launches, dev,
                                                   one_by_one=args.one_by_one != 0)}), flush=True)
        return
    if args.copy_calibration > 0:
        a = torch.empty(args.copy_calibration // 4, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    cfg = scenes.config_dam_break()
    st = cfg["state"]
    n = st["x"].shape[1]
    ctx = Context(cfg["dx"], cfg["n_cells"], n)
    ctx.set_materials(cfg["mats"])
    ctx.import_state({k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in st.items()})
    params = StepParams(cfg["dt"], cfg["gravity"], 1, cfg["pad"])
    wg = ctx.autotune(params, 4)
    if args.fused != "auto":
        ctx.set_fused(args.fused == "1")
        wg = (wg[0], 0 if ctx.fused() else 256, wg[2])
    ctx.substeps(params, args.warmup)
    ctx.sync()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.substeps(params, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ctx.sync()  # raises on a deferred device error (particle outside the lattice)
    # per-launch durations: the same number of substeps again, with HIP events around every launch
    # (outside the timed region: the events themselves cost the substep about 7 %)
    phase_ms = ctx.substeps_timed(params, args.steps)
    ctx.sync()
    copy_gbs = copy_bandwidth(dev)

    keys, data = ctx.grid()
    n_cells = int((data[:, 0, :] != 0).sum())
    n_solid, n_fluid = int(cfg["n_solid"]), int(cfg["n_fluid"])
    ab = algorithmic_bytes(n_solid, n_fluid, n_cells)
    fused = wg[1] == 0
    # the bytes of each phase's kernel (the fused pass in the G2P update phase)
    ab_phase = dict(ab, g2p_update=ab["g2p2g"]) if fused else ab
    # dominant kernel: the single launch with the longest average duration
    dom = max(("stress", "p2g", "grid", "g2p_move", "g2p_update"), key=lambda k: phase_ms[k])
    t_dom = phase_ms[dom] / 1e3 / args.steps
    achieved = ab_phase[dom] / t_dom / 1e9
    traffic = pmc_traffic("g2p2g" if (fused and dom == "g2p_update") else dom, cfg["name"])
    t_step = wall / args.steps
    out = {
        "metric": METRIC,
        "value": n * args.steps / wall,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "particles": n, "n_fluid": n_fluid, "n_solid": n_solid,
                   "dx": cfg["dx"], "dt": cfg["dt"], "active_nodes": n_cells,
                   "particle_blocks": ctx.num_particle_blocks(), "grid_blocks": len(keys),
                   "autotuned": {"g2p_move_wg": wg[0], "g2p_update_wg": wg[1], "sparse_p2g": wg[2],
                                 "fused_g2p_p2g": fused}},
        "roofline": {"bound": "hbm", "kernel": kernel_names(wg[0], wg[1])[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "algorithmic_bytes": ab_phase[dom], "avg_ms": t_dom * 1e3, "frac_of_copy": achieved / copy_gbs},
        "substep_roofline": {"achieved": ab["substep"] / t_step / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": ab["substep"] / t_step / 1e9 / HBM_PEAK_GBS,
                             "algorithmic_bytes": ab["substep"],
                             "frac_of_copy": ab["substep"] / t_step / 1e9 / copy_gbs},
        "copy_bandwidth": {"achieved": copy_gbs, "unit": "GB/s",
                           "method": "torch copy_ of 512 MiB device to device, (read + write) / time, 5 copies"},
        "phase_ms_per_step": {k: v / args.steps for k, v in phase_ms.items()},
        # every phase against the same two ceilings (algorithmic bytes / its event-timed duration)
        "kernel_roofline": {
            k: {"kernel": kernel_names(wg[0], wg[1])[k], "algorithmic_bytes": ab_phase[k],
                "avg_ms": phase_ms[k] / args.steps,
                "achieved_gbs": ab_phase[k] / (phase_ms[k] / 1e3 / args.steps) / 1e9 if phase_ms[k] > 0 else None,
                "frac": ab_phase[k] / (phase_ms[k] / 1e3 / args.steps) / 1e9 / HBM_PEAK_GBS if phase_ms[k] > 0 else None,
                "frac_of_copy": ab_phase[k] / (phase_ms[k] / 1e3 / args.steps) / 1e9 / copy_gbs
                if phase_ms[k] > 0 else None}
            for k in ("stress", "p2g", "grid", "g2p_move", "g2p_update")
            if not (fused and k in ("stress", "p2g"))},
    }
    ctx.close()
    if args.batched_launches > 0:
        out["batched"] = batched_rates(args.batched_scenes, args.batched_launches, dev)
    if args.adjoint:
        out["adjoint"] = adjoint_rates(dev)
    if args.sysid:
        out["sysid"] = sysid_rates()
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
