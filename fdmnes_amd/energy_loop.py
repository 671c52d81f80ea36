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
from .tau import split, symmetrize, tau_reduce

__all__ = ["EnergyLoop", "RepresentationLoop", "FdmCompLoop", "FP32", "FP64"]


class EnergyLoop:
    """Solves a cluster (``fdmnes_amd.cluster.Cluster``) at many energies.

    E_comp / basereel select calcMatRow's arithmetic branch (complex when
    E_comp or not basereel, mat.f90:96-104); atoms: [(first_row, nrows),
    ...] row blocks reduced to tau (default: none); fdm_comp drops the
    conjugate in the contraction.  The right-hand sides are the incoming
    waves calcMatRow builds (mat.f90:1380-1395, :1497-1507), and
    sigma(E) = -Im sum_L b_L^T x_L on them."""

    def __init__(self, cluster, precision: int = FP32, group: int = 12, E_comp: bool = True,
                 basereel: bool = False, atoms=None, fdm_comp: bool = False, sphere=None, tensor=None):
        """sphere: for a cluster with muffin-tin atoms, the
        ``fdmnes_amd.sphere.DeviceSphere`` of its atoms, which computes their
        phiato for every energy on the device (Sphere, sphere.f90:8); tau is
        then reduced over the atoms' row blocks unless atoms says otherwise.
        tensor: a ``fdmnes_amd.tensor.DeviceTensor`` of the absorbing atom
        (the sphere's radial integrals set by ``set_integrals``): each
        point's tau is then symmetrised as mat leaves it (mat.f90:596-641)
        and turned into the Cartesian cross-section tensors (Tenseur_car,
        tensor.f90:15)."""
        if not 1 <= group <= 32:
            raise ValueError("group: 1..32 energy points per grouped solve")
        self.cl = cluster
        self.G = int(group)
        self.E_comp, self.basereel, self.fdm_comp = bool(E_comp), bool(basereel), bool(fdm_comp)
        self.cal = self.E_comp or not self.basereel or bool(getattr(cluster, "ylm_comp", False))
        self.natome = int(getattr(cluster, "natome", 0))
        if bool(self.natome) != (sphere is not None):
            raise ValueError("a cluster with atoms needs its DeviceSphere (and one without atoms none)")
        self.sphere = sphere
        if tensor is not None and (sphere is None or getattr(sphere, "nrof", 0) == 0):
            raise ValueError("the tensors need the DeviceSphere's radial integrals (set_integrals)")
        self.tensor = tensor
        if atoms is None and self.natome:
            atoms = [(cluster.atom(a)["ianew"], cluster.atom(a)["nlmsa"]) for a in range(self.natome)]
        self.atoms = list(atoms or [])
        self.nso = cluster.nligne - cluster.nlmso
        self.asm = DeviceAssembler(cluster, max_points=min(self.G, 16))
        self.precision = precision
        self.sv = DeviceSolver(precision, self.cal, cluster.nlmso)
        lb1, lb2 = cluster.profile()
        if self.natome:
            zero = [np.zeros((cluster.atom(a)["nbord"], cluster.atom(a)["nlmsa"])) for a in range(self.natome)]
            rp, col, _, _ = cluster.assemble_atoms(np.zeros(cluster.npoint), 0.0, 0.0, 1.0, self.E_comp,
                                                   self.basereel, zero)
        else:
            rp, col, _, _ = cluster.assemble(np.zeros(cluster.npoint), 0.0, 0.0, 1.0, self.E_comp, self.basereel)
        self.slots = [self.sv.add_system(lb1, lb2, rp, col) for _ in range(self.G)]
        dt = torch.complex128 if self.cal else torch.float64
        n, nl = cluster.nligne, cluster.nlmso
        self.vals = [torch.zeros(cluster.nnz, dtype=dt, device="cuda") for _ in range(self.G)]
        self.rhs = [torch.zeros((n, nl), dtype=dt, device="cuda") for _ in range(self.G)]
        self.x = [torch.zeros((n, nl), dtype=dt, device="cuda") for _ in range(self.G)]
        self._filled = 0  # buffers holding an earlier group's assembly (update_only is valid for them)

    def run(self, V, energies, keep_solutions: bool = False, gradient: bool = False, energ=None) -> dict:
        """V: one potential (npoint, Ry) for every energy, or one per energy;
        energies: [(Enervide, Eimag, Ecinetic_out), ...] in Ry.  Returns
        {"sigma": (nE,), "tau": per energy a list of per-atom (nrows, nrows)
        complex arrays, "sweeps": refinement sweeps per energy, "x": the
        solutions (host arrays) when keep_solutions, and "dsigma_dV": (nE,
        npoint) when gradient: d sigma(E) / d V_i by the adjoint (the
        transposed substitutions with each group's factors, FP64 complex
        loops only).  With a tensor: "secdd", "secdq", "secqq" per energy
        (host arrays (ninitr, 3, ...), None for a multipole not computed);
        energ: per energy the photoelectron energy Energ of the quadrupole
        factor (default Enervide)."""
        if gradient and (self.precision != FP64 or not self.cal):
            raise ValueError("the gradient needs an FP64 loop in complex arithmetic")
        nE = len(energies)
        per_energy = isinstance(V, (list, tuple))
        if per_energy and len(V) != nE:
            raise ValueError("one potential per energy point")
        Vd = [torch.as_tensor(np.ascontiguousarray(v, np.float64)).cuda() for v in V] if per_energy \
            else torch.as_tensor(np.ascontiguousarray(V, np.float64)).cuda()
        sigma = np.zeros(nE)
        tau, sweeps, xs = [], [], []
        grad = np.zeros((nE, self.cl.npoint)) if gradient else None
        secs = {"secdd": [], "secdq": [], "secqq": []}
        energ = [e[0] for e in energies] if energ is None else list(energ)
        if len(energ) != nE:
            raise ValueError("one Energ per energy point")
        for e0 in range(0, nE, self.G):
            e1 = min(nE, e0 + self.G)
            k = e1 - e0
            vg = Vd[e0:e1] if per_energy else [Vd] * k
            # buffers 0..k-1 of a group after the first already hold this
            # structure's rows: only the energy-dependent entries change
            upd = self._filled >= k
            phi = rof = None
            if self.sphere is not None:  # the atoms' phiato of these energies (Ecinetic, Eimag, Enervide)
                ens = [(e[2], e[1], e[0]) for e in energies[e0:e1]]
                if self.tensor is not None:
                    phi, rof = self.sphere.run(ens, complex_out=self.cal, energ=energ[e0:e1])
                else:
                    phi = self.sphere.run(ens, complex_out=self.cal)
            self.asm.run(vg, energies[e0:e1], self.E_comp, self.basereel, self.vals[:k], self.rhs[:k],
                         update_only=upd, phi=phi)
            self._filled = max(self._filled, k)
            _, sw, _ = self.sv.solve(self.slots[:k], self.vals[:k], self.rhs[:k], self.x[:k])
            sweeps += list(sw)
            sigma[e0:e1] = self.sv.sigma(self.slots[:k], self.rhs[:k], self.x[:k])
            if gradient:  # the rows' values depend on V only through the diagonals
                g = [torch.zeros(self.cl.nligne, dtype=torch.float64, device="cuda") for _ in range(k)]
                self.sv.adjoint(self.slots[:k], self.rhs[:k], self.x[:k], g)
                for j in range(k):
                    grad[e0 + j] = g[j][:self.cl.npoint].cpu().numpy()
            if self.atoms:
                t, _, _ = tau_reduce(self.x[:k], self.nso, self.basereel, self.fdm_comp, self.atoms)
                if self.tensor is not None:
                    symmetrize(t, [na for _, na in self.atoms])
                    for name, o in zip(("secdd", "secdq", "secqq"), self.tensor.run(rof, t)):
                        secs[name] += [None] * k if o is None else [v.cpu().numpy() for v in o]
                tau += [[b.cpu().numpy() for b in split(ti, self.atoms)] for ti in t]
            if keep_solutions:
                xs += [xi.cpu().numpy() for xi in self.x[:k]]
        out = {"sigma": sigma, "tau": tau, "sweeps": sweeps}
        if self.tensor is not None:
            out.update(secs)
        if keep_solutions:
            out["x"] = xs
        if gradient:
            out["dsigma_dV"] = grad
        return out


class RepresentationLoop:
    """The energy loop of a cluster split into representations (the igrph
    loop, fdm.f90:3775, e.g. the eight systems of ``Cluster(..., irrep=g)``
    for the mirror group mmm): each representation's rows are assembled on
    the device, and every grouped solve takes as many energies of all the
    representations as fit in one launch of up to 32 systems, right-hand
    sides padded with zero columns to the largest count.  sigma(E) is
    returned per (energy, representation)."""

    MAX_SYSTEMS = 32

    def __init__(self, clusters, precision: int = FP32, E_comp: bool = True, basereel: bool = False):
        self.cls = list(clusters)
        if not self.cls or len(self.cls) > self.MAX_SYSTEMS:
            raise ValueError(f"1..{self.MAX_SYSTEMS} representations")
        self.E_comp, self.basereel = bool(E_comp), bool(basereel)
        self.cal = self.E_comp or not self.basereel or any(getattr(c, "ylm_comp", False) for c in self.cls)
        self.per = self.MAX_SYSTEMS // len(self.cls)  # energies per grouped solve
        self.nrhs = max(c.nlmso for c in self.cls)
        self.asm = [DeviceAssembler(c, max_points=min(self.per, 16)) for c in self.cls]
        self.sv = DeviceSolver(precision, self.cal, self.nrhs)
        dt = torch.complex128 if self.cal else torch.float64
        self.slots, self.vals, self.rhs, self.rhs_pad, self.x = [], [], [], [], []
        for c in self.cls:
            lb1, lb2 = c.profile()
            rp, col, _, _ = c.assemble(np.zeros(c.npoint), 0.0, 0.0, 1.0, self.E_comp, self.basereel)
            self.slots.append([self.sv.add_system(lb1, lb2, rp, col) for _ in range(self.per)])
            self.vals.append([torch.zeros(c.nnz, dtype=dt, device="cuda") for _ in range(self.per)])
            self.rhs.append([torch.zeros((c.nligne, c.nlmso), dtype=dt, device="cuda") for _ in range(self.per)])
            self.rhs_pad.append([torch.zeros((c.nligne, self.nrhs), dtype=dt, device="cuda") for _ in range(self.per)])
            self.x.append([torch.zeros((c.nligne, self.nrhs), dtype=dt, device="cuda") for _ in range(self.per)])
        self._filled = 0

    def run(self, V, energies, keep_solutions: bool = False) -> dict:
        """V: one potential per representation (its ``points``, Ry);
        energies: [(Enervide, Eimag, Ecinetic_out), ...] in Ry.  Returns
        {"sigma": (nE, nrep), "sweeps": (nE, nrep), and "x": per energy a
        list of per-representation solutions (nligne, nlmso) when
        keep_solutions}."""
        if len(V) != len(self.cls):
            raise ValueError("one potential per representation")
        Vd = [torch.as_tensor(np.ascontiguousarray(v, np.float64)).cuda() for v in V]
        nE, nr = len(energies), len(self.cls)
        sigma = np.zeros((nE, nr))
        sweeps = np.zeros((nE, nr), np.int64)
        xs = [[None] * nr for _ in range(nE)] if keep_solutions else None
        for e0 in range(0, nE, self.per):
            e1 = min(nE, e0 + self.per)
            k = e1 - e0
            upd = self._filled >= k
            slots, vals, rhs, x = [], [], [], []
            for r, c in enumerate(self.cls):
                self.asm[r].run([Vd[r]] * k, energies[e0:e1], self.E_comp, self.basereel, self.vals[r][:k],
                                self.rhs[r][:k], update_only=upd)
                for j in range(k):
                    self.rhs_pad[r][j][:, :c.nlmso].copy_(self.rhs[r][j])
                slots += self.slots[r][:k]
                vals += self.vals[r][:k]
                rhs += self.rhs_pad[r][:k]
                x += self.x[r][:k]
            self._filled = max(self._filled, k)
            _, sw, _ = self.sv.solve(slots, vals, rhs, x)
            sg = self.sv.sigma(slots, rhs, x)
            for r in range(nr):
                for j in range(k):
                    sigma[e0 + j, r] = sg[r * k + j]
                    sweeps[e0 + j, r] = sw[r * k + j]
                    if keep_solutions:
                        xs[e0 + j][r] = self.x[r][j][:, :self.cls[r].nlmso].cpu().numpy()
        out = {"sigma": sigma, "sweeps": sweeps}
        if keep_solutions:
            out["x"] = xs
        return out


class FdmCompLoop:
    """The energy loop of FDM_comp (mat.f90:135-295, Loop_RI): at complex
    energy each point solves the regular cluster system and the irregular
    one (``Cluster(..., atoms=..., irregular=True)``), and the atoms'
    amplitude products come from both solutions as Tau_RI.  Per group of
    energies: the device radial solver's regular and irregular phiato
    (``DeviceSphere.run_ri``), device assembly of both systems, one grouped
    solve of the 2 x group systems, Tau_RI on the device
    (``fdmnes_amd.tau.tau_ri``)."""

    def __init__(self, cluster, cluster_irr, sphere, precision: int = FP64, group: int = 8):
        if not 1 <= group <= 16:
            raise ValueError("group: 1..16 energy points (two systems each) per grouped solve")
        if not cluster.natome or not getattr(cluster_irr, "irregular", False):
            raise ValueError("a regular cluster with atoms and its irregular system")
        self.cl, self.ci, self.sphere, self.G = cluster, cluster_irr, sphere, int(group)
        self.asm = DeviceAssembler(cluster, max_points=self.G)
        self.asm_i = DeviceAssembler(cluster_irr, max_points=self.G)
        nl = cluster.nlmso
        self.sv = DeviceSolver(precision, True, nl)
        zero = [np.zeros((cluster.atom(a)["nbord"], cluster.atom(a)["nlmsa"])) for a in range(cluster.natome)]
        rp, col, _, _ = cluster.assemble_atoms(np.zeros(cluster.npoint), 0.0, 0.01, 1.0, True, False, zero)
        rpi, coli, _, _ = cluster_irr.assemble_atoms(np.zeros(cluster_irr.npoint), 0.0, 0.01, 1.0, True, False, zero,
                                                     zero)
        pr, pi = cluster.profile(), cluster_irr.profile()
        self.slots = [self.sv.add_system(*pr, rp, col) for _ in range(self.G)]
        self.slots_i = [self.sv.add_system(*pi, rpi, coli) for _ in range(self.G)]
        c128 = torch.complex128
        self.vals = [torch.zeros(cluster.nnz, dtype=c128, device="cuda") for _ in range(self.G)]
        self.vals_i = [torch.zeros(cluster_irr.nnz, dtype=c128, device="cuda") for _ in range(self.G)]
        self.rhs = [torch.zeros((cluster.nligne, nl), dtype=c128, device="cuda") for _ in range(self.G)]
        self.rhs_i = [torch.zeros((cluster_irr.nligne, nl), dtype=c128, device="cuda") for _ in range(self.G)]
        self.x = [torch.zeros_like(r) for r in self.rhs]
        self.x_i = [torch.zeros_like(r) for r in self.rhs_i]
        self.atoms = [(cluster.atom(a)["ianew"], cluster.atom(a)["nlmsa"]) for a in range(cluster.natome)]
        self.atoms_i = [(cluster_irr.atom(a)["ianew"], cluster_irr.atom(a)["nlmsa"]) for a in range(cluster.natome)]

    def run(self, V, energies) -> dict:
        """V: the potential (npoint, Ry); energies: [(Enervide, Eimag,
        Ecinetic_out), ...] with Eimag != 0.  Returns {"tau_ri": per energy
        a list of per-atom (2, nlmsa, nlmsa) complex arrays (i_RI = 1, 2)}."""
        from .tau import tau_ri
        if any(abs(e[1]) <= 1e-10 for e in energies):
            raise ValueError("FDM_comp solves at complex energy (Eimag != 0)")
        Vd = torch.as_tensor(np.ascontiguousarray(V, np.float64)).cuda()
        out = []
        for e0 in range(0, len(energies), self.G):
            ens = energies[e0:e0 + self.G]
            k = len(ens)
            pp = self.sphere.run_ri([(e[2], e[1], e[0]) for e in ens], complex_out=True)
            phi, phi2 = [p[0] for p in pp], [p[1] for p in pp]
            self.asm.run([Vd] * k, ens, True, False, self.vals[:k], self.rhs[:k], phi=phi)
            self.asm_i.run([Vd] * k, ens, True, False, self.vals_i[:k], self.rhs_i[:k], phi=phi, phi2=phi2)
            self.sv.solve(self.slots[:k] + self.slots_i[:k], self.vals[:k] + self.vals_i[:k],
                          self.rhs[:k] + self.rhs_i[:k], self.x[:k] + self.x_i[:k])
            t = tau_ri(self.x[:k], self.x_i[:k], self.atoms, self.atoms_i)
            for ti in t:
                flat, blocks, off = ti.cpu().numpy(), [], 0
                for _, na in self.atoms:
                    blocks.append(flat[off:off + 2 * na * na].reshape(2, na, na))
                    off += 2 * na * na
                out.append(blocks)
        return {"tau_ri": out}
