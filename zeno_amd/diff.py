"""Differentiable MPM rollouts on the GPU: checkpointed forward, reverse-mode
backward through the hand-written adjoint substep (zmpm_adjoint_substep), and a
torch.autograd.Function over the per-particle initial velocity and each
material's Young's modulus and Poisson ratio.

The forward keeps the resident state every `checkpoint_every` substeps.  The
backward walks the segments in reverse: it reloads a segment's checkpoint,
recomputes the segment's substeps once while saving the state at the start of
each (the recomputation is bit for bit the original, binning included), then
runs the adjoint substeps from the last to the first.  Adjoint arrays are
indexed by the caller's particle index, so they need no permutation when the
particles are re-binned.

The reference has no differentiable MPM; this mirrors its PyZpc harness
(ctypes calls with status codes, projects/PyZpc/Lib/zpy/utils.py:9-24).
"""
from __future__ import annotations

import numpy as np
import torch

from .zmpm import Context, StepParams


def lame(E, nu):
    """Lame parameters of FixedCorotated<float>{E, nu} (Generation.cu:41)."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def lame_grad(E, nu, mub, lamb):
    """(dL/dE, dL/dnu) from (dL/dmu, dL/dlambda)."""
    dmu_dE = 1.0 / (2.0 * (1.0 + nu))
    dmu_dnu = -E / (2.0 * (1.0 + nu) ** 2)
    dl_dE = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    dl_dnu = E * (1.0 + 2.0 * nu * nu) / ((1.0 + nu) ** 2 * (1.0 - 2.0 * nu) ** 2)
    return mub * dmu_dE + lamb * dl_dE, mub * dmu_dnu + lamb * dl_dnu


class Rollout:
    """A fixed scene (positions, materials other than E and nu, grid) rolled out for
    `nsteps` substeps from a given initial velocity and E, nu per material.

    state: numpy/torch SoA dict x (3,n), v (3,n) (ignored), C (9,n), F (9,n), mat (n,).
    mats: material dicts (model, rho, vol, bulk, scene); E and nu come from run().
    ctx: a Context (single or multi-scene) of capacity >= n.
    """

    def __init__(self, ctx: Context, state, mats, params, nsteps: int, checkpoint_every: int = 8):
        self.ctx = ctx
        self.dev = "cuda:%d" % ctx.device
        self.state = {k: torch.as_tensor(np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v).to(self.dev)
                      for k, v in state.items()}
        self.mats = [dict(m) for m in mats]
        self.params = params
        self.nsteps = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.n = int(self.state["x"].shape[1])
        self.ckpts = []
        self.E = self.nu = None

    def _segments(self):
        return [(s, min(self.k, self.nsteps - s)) for s in range(0, self.nsteps, self.k)]

    def run(self, v0: torch.Tensor, E, nu) -> torch.Tensor:
        """Forward rollout; returns the final positions (3, n) by caller index."""
        E = np.atleast_1d(np.asarray(E, np.float64))
        nu = np.atleast_1d(np.asarray(nu, np.float64))
        self.E, self.nu = E, nu
        mats = []
        for i, m in enumerate(self.mats):
            m = dict(m)
            m["E"], m["nu"] = float(E[i % len(E)]), float(nu[i % len(nu)])
            mats.append(m)
        self.ctx.set_materials(mats)
        st = dict(self.state)
        st["v"] = v0.detach().to(self.dev, torch.float32).contiguous()
        self.ctx.import_state(st)
        self.ckpts = []
        for _, length in self._segments():
            buf = self.ctx.new_state_buffer()
            self.ctx.state_save(buf)
            self.ckpts.append(buf)
            self.ctx.substeps(self.params, length)
        return self.ctx.export_state()["x"]

    def grad(self, xbar: torch.Tensor, vbar=None):
        """Reverse pass from dL/dx (and optionally dL/dv) of the final state.
        Returns dL/dv0 (3, n), dL/dE and dL/dnu per material (float64 numpy), and the
        adjoints of the initial x, C, F."""
        n, dev = self.n, self.dev
        adj = {"x": xbar.detach().to(dev, torch.float32).contiguous().clone(),
               "v": (vbar.detach().to(dev, torch.float32).contiguous().clone() if vbar is not None
                     else torch.zeros(3, n, device=dev)),
               "C": torch.zeros(9, n, device=dev), "F": torch.zeros(9, n, device=dev),
               "mu": torch.zeros(n, device=dev), "lam": torch.zeros(n, device=dev)}
        seg_states = [self.ctx.new_state_buffer() for _ in range(self.k)]
        for (start, length), ck in reversed(list(zip(self._segments(), self.ckpts))):
            self.ctx.state_load(ck)
            for j in range(length):
                self.ctx.state_save(seg_states[j])
                if j < length - 1:
                    self.ctx.substeps(self.params, 1)
            for j in reversed(range(length)):
                self.ctx.state_load(seg_states[j])
                self.ctx.adjoint_substep(self.params, adj)
        self.ctx.sync()
        mat = self.state["mat"].long()
        nm = len(self.mats)
        Eb = np.zeros(nm)
        nub = np.zeros(nm)
        for i in range(nm):
            sel = (mat == i).double()
            mub = float((adj["mu"].double() * sel).sum())
            lamb = float((adj["lam"].double() * sel).sum())
            Eb[i], nub[i] = lame_grad(self.E[i % len(self.E)], self.nu[i % len(self.nu)], mub, lamb)
        return {"v0": adj["v"], "x0": adj["x"], "C0": adj["C"], "F0": adj["F"], "E": Eb, "nu": nub}


class RolloutFunction(torch.autograd.Function):
    """x_final = rollout(v0, E, nu); E and nu are 1-D tensors (one entry per material)."""

    @staticmethod
    def forward(fctx, v0, E, nu, rollout: Rollout):
        fctx.rollout = rollout
        x = rollout.run(v0, E.detach().cpu().double().numpy(), nu.detach().cpu().double().numpy())
        return x

    @staticmethod
    def backward(fctx, xbar):
        g = fctx.rollout.grad(xbar)
        E_grad = torch.as_tensor(g["E"], dtype=torch.float64)
        nu_grad = torch.as_tensor(g["nu"], dtype=torch.float64)
        return g["v0"], E_grad, nu_grad, None


def rollout(v0, E, nu, sim: Rollout):
    return RolloutFunction.apply(v0, E, nu, sim)
