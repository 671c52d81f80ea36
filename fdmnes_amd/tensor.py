"""The absorbing atom's Cartesian cross-section tensors on the device
(``fdmgpu_tensor_*`` in include/fdmgpu.h): what the reference's Tenseur_car
(tensor.f90:15, with tens_ab :947) computes per energy point from the
atom's Taull and radial integrals, for the electric multipoles E1E1
(secdd), E1E2 (secdq) and E2E2 (secqq) of an FDM absorption calculation
(no TDDFT, optic, NRIXS, Green's-function or FDM_comp output, no singular
solution, no spin-orbit, one final-state spin, real harmonics).

The radial integrals come from ``DeviceSphere.run(..., energ=...)`` after
``DeviceSphere.set_integrals``; Taull from ``fdmnes_amd.tau.tau_reduce``
followed by ``fdmnes_amd.tau.symmetrize`` (as mat leaves it)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .batch import Ordered, library_stream

MAX_POINTS = 32
_V = C.c_void_p


class TensorTables(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("jseuil", "lseuil", "nbseuil", "core_resolved", "lmax", "lplus1", "lmoins1",
                                        "numat", "e1e1", "e1e2", "e2e2")] + \
               [(n, _V) for n in ("ldip", "lqua", "msymdd", "msymddi", "msymdq", "msymdqi", "msymqq", "msymqqi",
                                  "rot_int", "rot_atom_abs")] + \
               [(n, C.c_int) for n in ("rof_lmax", "rof_nbseuil", "rof_ip0", "rof_ip_max")] + \
               [("taull_offset", C.c_int64), ("taull_ld", C.c_int)]


def _sigs(L):
    if getattr(L, "_tens_sigs", False):
        return
    P = C.POINTER
    L.fdmgpu_tensor_create.argtypes = [P(TensorTables), C.c_int, _V, P(_V)]
    L.fdmgpu_tensor_destroy.argtypes = [_V]
    L.fdmgpu_tensor_destroy.restype = None
    L.fdmgpu_tensor_ninitr.argtypes = [_V]
    L.fdmgpu_tensor_run.argtypes = [_V, C.c_int, P(_V), P(_V), P(_V), P(_V), P(_V)]
    L._tens_sigs = True


def no_symmetry_tables():
    """Tensor shapes and symmetry tables of an absorber with no symmetry, not
    magnetic: every element allowed, the real part of each independent
    element its own class (numbered as the reference's tensdd / tensdq /
    tensqq number them for point group 1), imaginary parts excluded."""
    dd = np.zeros((3, 3), np.int32)
    k = 0
    for i in range(3):
        for j in range(i, 3):
            k += 1
            dd[i, j] = dd[j, i] = k
    dq = np.zeros((3, 3, 3), np.int32)
    k = 0
    for i in range(3):
        for j in range(3):
            for l in range(j, 3):
                k += 1
                dq[i, j, l] = dq[i, l, j] = k
    qq = np.zeros((3, 3, 3, 3), np.int32)
    k = 0
    for i in range(3):
        for j in range(i, 3):
            for a in range(3):
                for b in range(a, 3):
                    if qq[i, j, a, b]:
                        continue
                    k += 1
                    for (p, q) in ((i, j), (j, i)):
                        for (r, s) in ((a, b), (b, a)):
                            qq[p, q, r, s] = qq[r, s, p, q] = k
    z = np.zeros_like
    return dict(ldip=np.ones(3, np.int32), lqua=np.ones((3, 3), np.int32), msymdd=dd, msymddi=z(dd), msymdq=dq,
                msymdqi=z(dq), msymqq=qq, msymqqi=z(qq))


class DeviceTensor:
    """Cross-section tensors of many energy points on the GPU.

    Edge: jseuil, lseuil, nbseuil (coef_init's Jseuil, Lseuil, 1 or 2
    edges), core_resolved; lmax the probe's (nlm_probe = (lmax + 1)^2 <= the
    absorber's nlmsa); tables the tensor shapes and symmetry tables (dict of
    ldip (3,), lqua (3, 3), msymdd / msymddi (3, 3), msymdq / msymdqi (3, 3,
    3), msymqq / msymqqi (3, 3, 3, 3), the reference's index order);
    multipoles (E1E1, E1E2, E2E2); rof_layout: ``DeviceSphere.rof_layout``;
    taull_offset / taull_ld: where the absorber's (nlmsa, nlmsa) block sits in
    a point's flat tau and its nlmsa."""

    def __init__(self, jseuil: int, lseuil: int, nbseuil: int, lmax: int, tables, rof_layout: dict,
                 taull_offset: int, taull_ld: int, multipoles=(True, True, True), core_resolved: bool = False,
                 lplus1: bool = False, lmoins1: bool = False, numat: int = 26, rot_int=None, rot_atom_abs=None,
                 max_points: int = MAX_POINTS, stream: torch.cuda.Stream | None = None):
        L = _lib.lib()
        _sigs(L)
        self.L = L
        keep = {}

        def arr(name, a, dt, shape):
            a = np.ascontiguousarray(a, dt)
            if a.shape != shape:
                raise ValueError(f"{name}: shape {shape}")
            keep[name] = a
            return a.ctypes.data_as(_V)
        eye = np.eye(3)
        shapes = dict(ldip=(3,), lqua=(3, 3), msymdd=(3, 3), msymddi=(3, 3), msymdq=(3, 3, 3), msymdqi=(3, 3, 3),
                      msymqq=(3, 3, 3, 3), msymqqi=(3, 3, 3, 3))
        self.multipoles = tuple(bool(m) for m in multipoles)
        tb = TensorTables(int(jseuil), int(lseuil), int(nbseuil), int(bool(core_resolved)), int(lmax),
                          int(bool(lplus1)), int(bool(lmoins1)), int(numat), *[int(m) for m in self.multipoles],
                          *[arr(n, tables[n], np.int32, shp) for n, shp in shapes.items()],
                          arr("rot_int", eye if rot_int is None else rot_int, np.float64, (3, 3)),
                          arr("rot_atom_abs", eye if rot_atom_abs is None else rot_atom_abs, np.float64, (3, 3)),
                          int(rof_layout["lmax"]), int(rof_layout["nbseuil"]), int(rof_layout["ip0"]),
                          int(rof_layout["ip_max"]), int(taull_offset), int(taull_ld))
        self.stream = library_stream(stream)
        h = _V()
        _lib.check(L.fdmgpu_tensor_create(C.byref(tb), int(max_points), _V(self.stream.cuda_stream), C.byref(h)))
        self._h = h
        self.ninitr = int(L.fdmgpu_tensor_ninitr(h))
        self.nrof = int(rof_layout["nbseuil"]) * (int(rof_layout["ip_max"]) - int(rof_layout["ip0"]) + 1) * \
            (int(rof_layout["lmax"]) + 1)
        self.tau_need = int(taull_offset) + int(taull_ld) ** 2
        self.max_points = int(max_points)

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.fdmgpu_tensor_destroy(self._h)
            self._h = None

    def run(self, rof, tau):
        """rof, tau: per point the device radial integrals and flat tau.
        Returns (secdd, secdq, secqq): per point complex128 device tensors
        (ninitr, 3, 3), (ninitr, 3, 3, 3), (ninitr, 3, 3, 3, 3), or None for a
        multipole not computed."""
        n = len(rof)
        if len(tau) != n:
            raise ValueError("one tau per point")
        for t in list(rof) + list(tau):
            if t.dtype != torch.complex128 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("rof and tau: contiguous complex128 device tensors")
        if any(t.numel() < self.nrof for t in rof) or any(t.numel() < self.tau_need for t in tau):
            raise ValueError(f"rof: {self.nrof} values per point, tau: at least {self.tau_need}")
        shp = ((3, 3), (3, 3, 3), (3, 3, 3, 3))
        outs = [[torch.empty((self.ninitr,) + s, dtype=torch.complex128, device="cuda") for _ in range(n)]
                if m else None for m, s in zip(self.multipoles, shp)]

        def ptrs(ts):
            if ts is None:
                return None
            a = (_V * len(ts))()
            for j, t in enumerate(ts):
                a[j] = t.data_ptr()
            return a
        rc = 0
        with Ordered(self.stream):
            for k0 in range(0, n, self.max_points):
                k1 = min(n, k0 + self.max_points)
                rc = self.L.fdmgpu_tensor_run(self._h, k1 - k0, ptrs(rof[k0:k1]), ptrs(tau[k0:k1]),
                                              *[ptrs(o[k0:k1]) if o is not None else None for o in outs])
                if rc:
                    break
        _lib.check(rc)
        return tuple(outs)
