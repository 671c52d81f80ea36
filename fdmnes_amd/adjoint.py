"""d sigma(E) / d V through the adjoint solve, behind a torch.autograd.Function
(for fitting potential parameters; the reference has no derivatives).

sigma_E(V) = -Im sum_L s_L^T x_L with A(E, V) x_L = s_L.  The forward pass
solves the energy points in groups of ``slots`` on the device (FP64); the
factors of the last group stay resident.  The backward pass runs the
transposed solves (A^T lam = s) with resident factors, and refactorises the
points whose factors were evicted (checkpoint / recompute)."""
from __future__ import annotations

import numpy as np
import torch

from . import batch as B
from . import configs as cfg


class SigmaEngine:
    """One cluster, its sources and a device solver with ``slots`` slots."""

    def __init__(self, cluster, sources: np.ndarray, eimag: float, slots: int = 4):
        self.cl = cluster
        self.lb1, self.lb2 = cluster.profile()
        self.eimag = eimag
        self.src = torch.from_numpy(np.ascontiguousarray(sources, np.complex128)).cuda()
        self.sv = B.DeviceSolver(B.FP64, True, sources.shape[1])
        self.rp = self.col = None
        self.G = slots
        self._structure = None
        self.resident = {}  # energy index -> slot holding its factors
        self.recomputed = 0

    def _vals(self, V: np.ndarray, E: float):
        rp, col, vals, _ = self.cl.assemble(V, E, self.eimag, cfg.ecinetic_out(E), True, False)
        if self._structure is None:
            self._structure = (rp, col)
            self.slots = [self.sv.add_system(self.lb1, self.lb2, rp, col) for _ in range(self.G)]
        return torch.from_numpy(vals.astype(np.complex128)).cuda()

    def forward(self, V: np.ndarray, energies) -> tuple[np.ndarray, list]:
        sig = np.zeros(len(energies))
        xs = []
        self.resident = {}
        for g0 in range(0, len(energies), self.G):
            idx = list(range(g0, min(g0 + self.G, len(energies))))
            vals = [self._vals(V, float(energies[i])) for i in idx]
            x = [torch.zeros_like(self.src) for _ in idx]
            sl = self.slots[: len(idx)]
            self.sv.solve(sl, vals, [self.src] * len(idx), x)
            sig[idx] = self.sv.sigma(sl, [self.src] * len(idx), x)
            xs += x
            self.resident = {i: s for i, s in zip(idx, sl)}
        return sig, xs

    def gradient(self, V: np.ndarray, energies, xs, weights: np.ndarray) -> np.ndarray:
        """sum_E weights[E] d sigma_E / d V (per interior point)."""
        n = self.cl.nligne
        total = torch.zeros(n, dtype=torch.float64, device="cuda")
        order = sorted(range(len(energies)), key=lambda i: (i not in self.resident, i))
        for g0 in range(0, len(order), self.G):
            idx = order[g0:g0 + self.G]
            if all(i in self.resident for i in idx):
                sl = [self.resident[i] for i in idx]
            else:  # factors evicted: refactorise these points (recompute)
                sl = self.slots[: len(idx)]
                vals = [self._vals(V, float(energies[i])) for i in idx]
                x = [torch.zeros_like(self.src) for _ in idx]
                self.sv.solve(sl, vals, [self.src] * len(idx), x)
                self.recomputed += len(idx)
                self.resident = {i: s for i, s in zip(idx, sl)}
            grads = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in idx]
            self.sv.adjoint(sl, [self.src] * len(idx), [xs[i] for i in idx], grads)
            for i, gr in zip(idx, grads):
                total += float(weights[i]) * gr
        return total[: self.cl.npoint].cpu().numpy()


class SigmaFunction(torch.autograd.Function):
    """sigma(E) for a batch of energies as a differentiable function of V."""

    @staticmethod
    def forward(ctx, V: torch.Tensor, engine: SigmaEngine, energies):
        Vn = V.detach().cpu().numpy().astype(np.float64)
        sig, xs = engine.forward(Vn, energies)
        ctx.engine, ctx.energies, ctx.xs, ctx.Vn = engine, energies, xs, Vn
        return torch.from_numpy(sig).to(V.device)

    @staticmethod
    def backward(ctx, grad_sigma: torch.Tensor):
        w = grad_sigma.detach().cpu().numpy().astype(np.float64)
        g = ctx.engine.gradient(ctx.Vn, ctx.energies, ctx.xs, w)
        return torch.from_numpy(g).to(grad_sigma.device), None, None
