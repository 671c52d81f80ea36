"""The atoms' radial solutions and their values at the atoms' border points,
phiato, on the device (``fdmgpu_sphere_*`` in include/fdmgpu.h): the work of
Sphere (sphere.f90:8-210) per energy point in the reference's energy loop
(fdm.f90:3733) for atoms with a spherical potential, one spin, the regular
solution.  The output is the device assembler's phiato layout, so it feeds
``fdmnes_amd.assemble.DeviceAssembler.run(..., phi=...)`` directly."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .batch import Ordered, library_stream

MAX_POINTS = 16
_V = C.c_void_p


class SphereTables(C.Structure):
    _fields_ = [("natome", C.c_int), ("nr", _V), ("r", _V), ("V", _V), ("Z", _V), ("lmax", _V), ("rmtg", _V),
                ("nbord", _V), ("bxyz", _V), ("pos", _V), ("ylmato", _V), ("relativiste", C.c_int),
                ("V_intmax", C.c_double)]


def _sigs(L):
    if getattr(L, "_sph_sigs", False):
        return
    P = C.POINTER
    L.fdmgpu_sphere_create.argtypes = [P(SphereTables), C.c_int, _V, P(_V)]
    L.fdmgpu_sphere_destroy.argtypes = [_V]
    L.fdmgpu_sphere_destroy.restype = None
    L.fdmgpu_sphere_run.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_int, P(_V), P(_V)]
    L.fdmgpu_sphere_run_ri.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_int, P(_V),
                                       P(_V), P(_V)]
    L.fdmgpu_sphere_set_integrals.argtypes = [_V, C.c_int, C.c_int, P(C.c_double), P(C.c_double), C.c_double,
                                              C.c_int, C.c_int]
    L.fdmgpu_sphere_run_integrals.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                                              P(C.c_double), C.c_int, P(_V), P(_V), P(_V)]
    L._sph_sigs = True


class DeviceSphere:
    """Radial solutions of a cluster's atoms for many energies on the GPU.

    atoms: per atom a dict with r (radial grid, bohr), V (spherical
    potential on it, Ry), Z, rmtg (matching radius, bohr), lmax, bxyz
    (nbord, 3) the atom's border points and pos (3,) the atom (bohr), ylmato
    (nbord, (lmax+1)^2) the real harmonics about the atom at them.  For a
    ``fdmnes_amd.cluster.Cluster`` with atoms, :func:`atoms_of` builds the
    border points and harmonics from the cluster."""

    def __init__(self, atoms, relativiste: bool = False, V_intmax: float = 1.0e6, max_points: int = MAX_POINTS,
                 stream: torch.cuda.Stream | None = None):
        L = _lib.lib()
        _sigs(L)
        self.L = L
        self.natome = len(atoms)
        self.nlmsa = [(int(a["lmax"]) + 1) ** 2 for a in atoms]
        self.lmax = [int(a["lmax"]) for a in atoms]
        self.nbord = [len(a["bxyz"]) for a in atoms]
        self.nphi = sum(nb * nl for nb, nl in zip(self.nbord, self.nlmsa))
        self.ntau = sum(int(a["lmax"]) + 1 for a in atoms)
        for a, nb, nl in zip(atoms, self.nbord, self.nlmsa):
            if np.shape(a["ylmato"]) != (nb, nl) or len(a["r"]) != len(a["V"]):
                raise ValueError("each atom: ylmato (nbord, (lmax+1)^2), V on the radial grid")
        k = dict(nr=np.array([len(a["r"]) for a in atoms], np.int32),
                 r=np.concatenate([np.asarray(a["r"], np.float64) for a in atoms]),
                 V=np.concatenate([np.asarray(a["V"], np.float64) for a in atoms]),
                 Z=np.array([int(a["Z"]) for a in atoms], np.int32),
                 lmax=np.array([int(a["lmax"]) for a in atoms], np.int32),
                 rmtg=np.array([float(a["rmtg"]) for a in atoms], np.float64),
                 nbord=np.array(self.nbord, np.int32),
                 bxyz=np.concatenate([np.asarray(a["bxyz"], np.float64).reshape(-1) for a in atoms]),
                 pos=np.concatenate([np.asarray(a["pos"], np.float64).reshape(3) for a in atoms]),
                 ylmato=np.concatenate([np.asarray(a["ylmato"], np.float64).reshape(-1) for a in atoms]))
        self._keep = {n: np.ascontiguousarray(v) for n, v in k.items()}
        names = ("nr", "r", "V", "Z", "lmax", "rmtg", "nbord", "bxyz", "pos", "ylmato")
        tb = SphereTables(self.natome, *[self._keep[n].ctypes.data_as(_V) for n in names], int(bool(relativiste)),
                          float(V_intmax))
        self.stream = library_stream(stream)
        h = _V()
        _lib.check(L.fdmgpu_sphere_create(C.byref(tb), int(max_points), _V(self.stream.cuda_stream), C.byref(h)))
        self._h = h
        self.max_points = int(max_points)

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.fdmgpu_sphere_destroy(self._h)
            self._h = None

    def set_integrals(self, atom: int, psii, Eseuil, Rmtsd: float, ip0: int = 1, ip_max: int = 2):
        """Names the absorbing atom (0-based) for the radial integrals of
        the golden rule (Radial_integral, sphere.f90:1490): its core
        functions psii (nbseuil, nr of the atom), the edge energies Eseuil
        (Ry), the integration radius Rmtsd (bohr) and the powers ip0..ip_max
        of r (1 dipole, 2 quadrupole)."""
        psii = np.ascontiguousarray(np.atleast_2d(psii), np.float64)
        es = np.ascontiguousarray(np.atleast_1d(Eseuil), np.float64)
        D = C.POINTER(C.c_double)
        if len(es) != psii.shape[0]:
            raise ValueError("one edge energy per core function")
        _lib.check(self.L.fdmgpu_sphere_set_integrals(self._h, int(atom), psii.shape[0], psii.ctypes.data_as(D),
                                                      es.ctypes.data_as(D), float(Rmtsd), int(ip0), int(ip_max)))
        self.rof_layout = dict(atom=int(atom), nbseuil=psii.shape[0], ip0=int(ip0), ip_max=int(ip_max),
                               lmax=int(self.lmax[atom]))
        self.nrof = psii.shape[0] * (ip_max - ip0 + 1) * (self.lmax[atom] + 1)

    def run_ri(self, energies, complex_out: bool = True):
        """The regular and the irregular solutions' phiato (Sch_radial_solsing,
        sphere.f90:2591): per point two device tensors (phi, phi2) in the
        layout of ``run``, for the irregular cluster system."""
        n = len(energies)
        dt = torch.complex128 if complex_out else torch.float64
        phi = [torch.empty(self.nphi, dtype=dt, device="cuda") for _ in range(n)]
        phi2 = [torch.empty(self.nphi, dtype=dt, device="cuda") for _ in range(n)]
        E = np.array(energies, np.float64).reshape(n, 3)
        D = C.POINTER(C.c_double)

        def ptrs(ts):
            arr = (_V * len(ts))()
            for j, t in enumerate(ts):
                arr[j] = t.data_ptr()
            return arr
        rc = 0
        with Ordered(self.stream):
            for k0 in range(0, n, self.max_points):
                k1 = min(n, k0 + self.max_points)
                cols = [np.ascontiguousarray(E[k0:k1, j]) for j in range(3)]
                rc = self.L.fdmgpu_sphere_run_ri(self._h, k1 - k0, *[c.ctypes.data_as(D) for c in cols],
                                                 int(bool(complex_out)), ptrs(phi[k0:k1]), ptrs(phi2[k0:k1]), None)
                if rc:
                    break
        _lib.check(rc)
        return list(zip(phi, phi2))

    def run(self, energies, complex_out: bool = True, with_tau: bool = False, energ=None):
        """energies: list of (Ecinetic, Eimag, Enervide) in Ry.  Returns per
        point a device tensor of the nphi phiato values (complex128, or
        float64 with complex_out False, which needs real radial solutions)
        and, with with_tau, per point the sum (lmax+1) complex Tau.  With
        energ (per point the photoelectron energy Energ, Ry) the absorbing
        atom's radial integrals come last in the returned tuple: per point
        a complex128 device tensor (nbseuil, ip_max - ip0 + 1, lmax + 1)."""
        n = len(energies)
        dt = torch.complex128 if complex_out else torch.float64
        phi = [torch.empty(self.nphi, dtype=dt, device="cuda") for _ in range(n)]
        tau = [torch.empty(self.ntau, dtype=torch.complex128, device="cuda") for _ in range(n)] if with_tau else None
        E = np.array(energies, np.float64).reshape(n, 3)
        D = C.POINTER(C.c_double)
        rof = None
        if energ is not None:
            if getattr(self, "nrof", 0) == 0:
                raise ValueError("set_integrals first")
            en = np.ascontiguousarray(np.atleast_1d(energ), np.float64)
            if len(en) != n:
                raise ValueError("one Energ per point")
            lay = self.rof_layout
            rof = [torch.empty((lay["nbseuil"], lay["ip_max"] - lay["ip0"] + 1, lay["lmax"] + 1),
                               dtype=torch.complex128, device="cuda") for _ in range(n)]

        def ptrs(ts):
            arr = (_V * len(ts))()
            for j, t in enumerate(ts):
                arr[j] = t.data_ptr()
            return arr
        rc = 0
        with Ordered(self.stream):
            for k0 in range(0, n, self.max_points):
                k1 = min(n, k0 + self.max_points)
                cols = [np.ascontiguousarray(E[k0:k1, j]) for j in range(3)]
                if rof is None:
                    rc = self.L.fdmgpu_sphere_run(self._h, k1 - k0, *[c.ctypes.data_as(D) for c in cols],
                                                  int(bool(complex_out)), ptrs(phi[k0:k1]),
                                                  ptrs(tau[k0:k1]) if tau is not None else None)
                else:
                    ek = np.ascontiguousarray(en[k0:k1])
                    rc = self.L.fdmgpu_sphere_run_integrals(self._h, k1 - k0, *[c.ctypes.data_as(D) for c in cols],
                                                            ek.ctypes.data_as(D), int(bool(complex_out)),
                                                            ptrs(phi[k0:k1]),
                                                            ptrs(tau[k0:k1]) if tau is not None else None,
                                                            ptrs(rof[k0:k1]))
                if rc:
                    break
        _lib.check(rc)
        out = (phi, tau) if with_tau else (phi,)
        if rof is not None:
            out = out + (rof,)
        return out if len(out) > 1 else out[0]


def atoms_of(cluster, r_grids, potentials, Z, rmtg, posi):
    """Sphere's per-atom inputs for a ``Cluster`` with atoms: the radial
    grids, potentials, atomic numbers and matching radii given, the border
    points (all nbord of each atom) and their harmonics from the cluster."""
    pts = cluster.points
    out = []
    for a in range(cluster.natome):
        d = cluster.atom(a)
        if np.any(d["ibord"] >= len(pts)):
            raise ValueError("an atom's border point lies outside the cluster's interior points")
        xyz = pts[d["ibord"]]
        lmax = int(round(np.sqrt(d["nlmsa"]))) - 1
        out.append(dict(r=r_grids[a], V=potentials[a], Z=Z[a], rmtg=rmtg[a], lmax=lmax, bxyz=xyz,
                        pos=np.asarray(posi[a], np.float64), ylmato=d["ylmato"]))
    return out
