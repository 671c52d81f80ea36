"""The energy loop of one cluster structure on the GPU (fdm.f90:3493-4192
calls mat, mat.f90:20, once per energy point, serially): energies are taken
in groups of up to ``group`` points; for each group the rows are assembled
on the device (only the energy-dependent entries after the first group),
the systems are solved in one grouped launch sequence (FP64, or FP32
factors with FP64 refinement), and each solution is reduced on the device
to sigma(E) and to the atom amplitude products tau (mat.f90:297-405).  Only
sigma and tau (and, on request, the solutions) come back to the host."""
from __future__ import annotations

import numpy as np
import torch

from .assemble import DeviceAssembler
from .batch import FP32, FP64, DeviceSolver
from .tau import split, tau_reduce

__all__ = ["EnergyLoop", "FP32", "FP64"]


class EnergyLoop:
    """Solves a cluster (``fdmnes_amd.cluster.Cluster``) at many energies.

    E_comp / basereel select calcMatRow's arithmetic branch (complex when
    E_comp or not basereel, mat.f90:96-104); atoms: [(first_row, nrows),
    ...] row blocks reduced to tau (default: none); fdm_comp drops the
    conjugate in the contraction.  The right-hand sides are the incoming
    waves calcMatRow builds (mat.f90:1380-1395, :1497-1507), and
    sigma(E) = -Im sum_L b_L^T x_L on them."""

    def __init__(self, cluster, precision: int = FP32, group: int = 12, E_comp: bool = True,
                 basereel: bool = False, atoms=None, fdm_comp: bool = False):
        if not 1 <= group <= 16:
            raise ValueError("group: 1..16 energy points per grouped solve")
        self.cl = cluster
        self.G = int(group)
        self.E_comp, self.basereel, self.fdm_comp = bool(E_comp), bool(basereel), bool(fdm_comp)
        self.cal = self.E_comp or not self.basereel
        self.atoms = list(atoms or [])
        self.nso = cluster.nligne - cluster.nlmso
        self.asm = DeviceAssembler(cluster, max_points=self.G)
        self.sv = DeviceSolver(precision, self.cal, cluster.nlmso)
        lb1, lb2 = cluster.profile()
        rp, col, _, _ = cluster.assemble(np.zeros(cluster.npoint), 0.0, 0.0, 1.0, self.E_comp, self.basereel)
        self.slots = [self.sv.add_system(lb1, lb2, rp, col) for _ in range(self.G)]
        dt = torch.complex128 if self.cal else torch.float64
        n, nl = cluster.nligne, cluster.nlmso
        self.vals = [torch.zeros(cluster.nnz, dtype=dt, device="cuda") for _ in range(self.G)]
        self.rhs = [torch.zeros((n, nl), dtype=dt, device="cuda") for _ in range(self.G)]
        self.x = [torch.zeros((n, nl), dtype=dt, device="cuda") for _ in range(self.G)]
        self._filled = 0  # buffers holding an earlier group's assembly (update_only is valid for them)

    def run(self, V, energies, keep_solutions: bool = False) -> dict:
        """V: one potential (npoint, Ry) for every energy, or one per energy;
        energies: [(Enervide, Eimag, Ecinetic_out), ...] in Ry.  Returns
        {"sigma": (nE,), "tau": per energy a list of per-atom (nrows, nrows)
        complex arrays, "sweeps": refinement sweeps per energy, and
        "x": the solutions (host arrays) when keep_solutions}."""
        nE = len(energies)
        per_energy = isinstance(V, (list, tuple))
        if per_energy and len(V) != nE:
            raise ValueError("one potential per energy point")
        Vd = [torch.as_tensor(np.ascontiguousarray(v, np.float64)).cuda() for v in V] if per_energy \
            else torch.as_tensor(np.ascontiguousarray(V, np.float64)).cuda()
        sigma = np.zeros(nE)
        tau, sweeps, xs = [], [], []
        for e0 in range(0, nE, self.G):
            e1 = min(nE, e0 + self.G)
            k = e1 - e0
            vg = Vd[e0:e1] if per_energy else [Vd] * k
            # buffers 0..k-1 of a group after the first already hold this
            # structure's rows: only the energy-dependent entries change
            upd = self._filled >= k
            self.asm.run(vg, energies[e0:e1], self.E_comp, self.basereel, self.vals[:k], self.rhs[:k],
                         update_only=upd)
            self._filled = max(self._filled, k)
            _, sw, _ = self.sv.solve(self.slots[:k], self.vals[:k], self.rhs[:k], self.x[:k])
            sweeps += list(sw)
            sigma[e0:e1] = self.sv.sigma(self.slots[:k], self.rhs[:k], self.x[:k])
            if self.atoms:
                t, _, _ = tau_reduce(self.x[:k], self.nso, self.basereel, self.fdm_comp, self.atoms)
                tau += [[b.cpu().numpy() for b in split(ti, self.atoms)] for ti in t]
            if keep_solutions:
                xs += [xi.cpu().numpy() for xi in self.x[:k]]
        out = {"sigma": sigma, "tau": tau, "sweeps": sweeps}
        if keep_solutions:
            out["x"] = xs
        return out
