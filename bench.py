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


def algorithmic_bytes(n_solid, n_fluid, n_cells):
    """Per-phase algorithmic HBM bytes of one substep (FP32 state, int32 material id).

    Particle state: x3 v3 C9 + F9 (fixed corotated) or J1 (J-fluid), + material id.
      P2G reads x, v, C, F|J, id                -> 100 B solid, 68 B fluid; writes grid m, mv: 16 B / active node
      grid update reads m, mv, writes v         -> 28 B / active node
      G2P reads x, F|J, id (+ grid v 12 B/node) -> 52 B solid, 20 B fluid; writes x, v, C, F|J: 96 B / 64 B
    Whole substep (fused P2G/G2P ideal): 196 B solid + 132 B fluid + 32 B per active node.
    """
    return {
        "p2g": 100 * n_solid + 68 * n_fluid + 16 * n_cells,
        "grid": 28 * n_cells,
        "g2p_update": (52 + 96) * n_solid + (20 + 64) * n_fluid + 12 * n_cells,
        "g2p_move": 12 * (n_solid + n_fluid) + 12 * n_cells,
        "substep": 196 * n_solid + 132 * n_fluid + 32 * n_cells,
    }


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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed substeps before timing")
    ap.add_argument("--cpu-steps", type=int, default=100, help="substeps of the CPU baseline sample (0 = skip)")
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
    ctx.substeps(params, args.warmup)
    ctx.sync()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phase_ms = ctx.substeps_timed(params, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ctx.sync()  # raises on a deferred device error (particle outside the lattice)

    keys, data = ctx.grid()
    n_cells = int((data[:, 0, :] != 0).sum())
    n_solid, n_fluid = int(cfg["n_solid"]), int(cfg["n_fluid"])
    ab = algorithmic_bytes(n_solid, n_fluid, n_cells)
    dom = max(("p2g", "grid", "g2p_move", "g2p_update"), key=lambda k: phase_ms[k])
    t_dom = phase_ms[dom] / 1e3 / args.steps
    achieved = ab[dom] / t_dom / 1e9
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
                   "particle_blocks": ctx.num_particle_blocks(), "grid_blocks": len(keys)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": ab[dom], "avg_ms": t_dom * 1e3},
        "substep_roofline": {"achieved": ab["substep"] / t_step / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": ab["substep"] / t_step / 1e9 / HBM_PEAK_GBS,
                             "algorithmic_bytes": ab["substep"]},
        "phase_ms_per_step": {k: v / args.steps for k, v in phase_ms.items()},
    }
    ctx.close()
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
