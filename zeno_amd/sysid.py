"""Batched differentiable system identification on the GPU.

Many scenes run in one multi-scene context; each scene has its own hidden
Young's modulus and Poisson ratio.  Target final positions come from a
rollout with the hidden parameters; Adam then fits (ln E, nu) of every scene
at once from the adjoint gradients of L = sum over scenes of the mean squared
distance between the scene's final positions and its targets.

Workload (the identification configuration): scenes of 16^3 cells x 8
particles per cell (32,768 particles, side 1/6) at dx = 1/96, centred;
hidden E log-uniform in [1e4, 1e6], nu uniform in [0.15, 0.4] (SplitMix64,
seed 7 + scene); 48 substeps at dt = 1e-4.  Every scene starts with the same
velocity field, a uniform drift plus a stretch along x and a squeeze along y
(rates 30/s and -20/s): without an initial deformation a free-falling block
keeps F = I and E, nu would not affect its motion.  Adam: lr 0.05 on ln E,
0.01 on nu, from E = 1e5, nu = 0.3.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import scenes
from .diff import Rollout
from .zmpm import BOUNDARY_STICKY_BOX, Context, MODEL_FCR, StepParams


def make_scenes(n_scenes=64, cells=16, ppc=8, n_cells=96, seed=7, dt=1e-4):
    dx = 1.0 / n_cells
    lo = (n_cells - cells) // 2
    out = []
    for s in range(n_scenes):
        u = scenes.uniform24(seed + s, np.arange(2, dtype=np.uint64))
        E = float(math.exp(math.log(1e4) + u[0] * (math.log(1e6) - math.log(1e4))))
        nu = float(0.15 + 0.25 * u[1])
        st, mats = scenes.elastic_box(dx, (lo, lo, lo), (lo + cells,) * 3, ppc, seed=seed + 1000 + s, E=E, nu=nu)
        c = st["x"].mean(axis=1, keepdims=True)
        st["v"] = (np.array([0.2, -0.3, 0.1])[:, None]
                   + np.array([30.0, -20.0, 0.0])[:, None] * (st["x"] - c)).astype(np.float32)
        out.append(dict(name=f"sysid_{s}", state=st, mats=mats, dx=dx, n_cells=n_cells, dt=dt,
                        gravity=(0.0, -9.8, 0.0), pad=3, E=E, nu=nu))
    return out


def identify(batch, steps=48, iters=50, lr_lnE=0.05, lr_nu=0.01, E0=1e5, nu0=0.3, checkpoint_every=8,
             log=None, keep_target=False):
    """Fit (ln E, nu) per scene; returns the loss curve, the fitted and hidden parameters and timings.

    Also returned: the first iteration's per-scene gradient (dL/dE, dL/dnu) at (E0, nu0) as
    "first_grad" (2, S), and with keep_target the target final positions (3, n) on the host
    and each scene's particle range, so the gradient can be checked against finite
    differences of another simulator."""
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    n = st["x"].shape[1]
    ctx = Context.scenes(descs, n)
    sim = Rollout(ctx, st, mats, None, steps, checkpoint_every=checkpoint_every)
    dev = sim.dev
    v0 = torch.from_numpy(np.ascontiguousarray(st["v"])).to(dev)
    hidden_E = np.array([b["E"] for b in batch])
    hidden_nu = np.array([b["nu"] for b in batch])
    target = sim.run(v0, hidden_E, hidden_nu).clone()
    # per particle weight 1 / n_scene: the loss is the sum over scenes of each scene's mean
    w = torch.empty(n, device=dev)
    for a, b in ranges:
        w[a:b] = 1.0 / (b - a)
    theta = np.stack([np.full(len(batch), math.log(E0)), np.full(len(batch), nu0)])  # (2, S)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    lr = np.array([[lr_lnE], [lr_nu]])
    b1, b2, eps = 0.9, 0.999, 1e-8
    curve = []
    first_grad = None
    t0 = time.perf_counter()
    for it in range(1, iters + 1):
        E, nu = np.exp(theta[0]), theta[1]
        x = sim.run(v0, E, nu)
        d = x - target
        loss = float((w * (d * d).sum(dim=0)).sum())
        curve.append(loss)
        g = sim.grad((2.0 * w)[None, :] * d)
        if first_grad is None:
            first_grad = np.stack([np.asarray(g["E"], np.float64), np.asarray(g["nu"], np.float64)])
        grad = np.stack([g["E"] * E, g["nu"]])  # d/d ln E = E d/dE
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad * grad
        mh, vh = m / (1 - b1 ** it), v / (1 - b2 ** it)
        theta = theta - lr * mh / (np.sqrt(vh) + eps)
        theta[1] = np.clip(theta[1], 0.01, 0.49)
        if log:
            log(it, loss)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    E, nu = np.exp(theta[0]), theta[1]
    out = {"loss": curve, "E": E, "nu": nu, "hidden_E": hidden_E, "hidden_nu": hidden_nu,
           "E_rel_err": np.abs(E - hidden_E) / hidden_E, "nu_rel_err": np.abs(nu - hidden_nu) / hidden_nu,
           "seconds": elapsed, "particles": n, "scenes": len(batch), "steps": steps, "iters": iters,
           "first_grad": first_grad, "ranges": ranges}
    if keep_target:
        out["target"] = target.cpu().numpy()
    ctx.close()
    return out
