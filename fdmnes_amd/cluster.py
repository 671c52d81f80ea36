"""Synthetic FDM clusters (grid + stencil + outer sphere) and their
per-energy systems, through libfdmgpu's cluster builder
(include/fdmgpu_cluster.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

BOHR = 0.52917721067  # Angstrom, main.f90:59
RYDB = 13.605693012905  # eV per Rydberg from main.f90:64's CODATA-2014 constants


def _L():
    L = _lib.load_host()
    if not hasattr(L, "_cluster_sigs"):
        L.fdmgpu_cluster_create.restype = C.c_int
        L.fdmgpu_cluster_create.argtypes = [C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_destroy.argtypes = [C.c_void_p]
        L.fdmgpu_cluster_sizes.argtypes = [C.c_void_p, _lib.ip]
        L.fdmgpu_cluster_points.argtypes = [C.c_void_p, _lib.dp]
        L.fdmgpu_cluster_profile.argtypes = [C.c_void_p, _lib.ip, _lib.ip]
        L.fdmgpu_cluster_nnz.restype = C.c_int64
        L.fdmgpu_cluster_nnz.argtypes = [C.c_void_p]
        L.fdmgpu_cluster_assemble.restype = C.c_int
        L.fdmgpu_cluster_assemble.argtypes = [C.c_void_p, _lib.dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                              C.c_int, _lib.lp, _lib.ip, _lib.dp, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L._cluster_sigs = True
    return L


class Cluster:
    """A spherical cluster of radius ``rsort`` and grid step ``adimp``
    (bohr) with ``(lmaxso+1)**2`` outer-sphere unknowns."""

    def __init__(self, rsort: float, adimp: float, lmaxso: int):
        L = _L()
        h = C.c_void_p()
        rc = L.fdmgpu_cluster_create(rsort, adimp, lmaxso, C.byref(h))
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_create failed ({rc})")
        self._h = h
        s = np.zeros(6, np.int32)
        L.fdmgpu_cluster_sizes(h, s)
        self.npoint, self.npso, self.nligne, self.nlmso, self.nsortf, self.nsort = map(int, s)
        self.lmaxso = lmaxso
        self.rsort, self.adimp = rsort, adimp
        self.nnz = int(L.fdmgpu_cluster_nnz(h))

    def __del__(self):
        if getattr(self, "_h", None):
            _L().fdmgpu_cluster_destroy(self._h)
            self._h = None

    @property
    def points(self) -> np.ndarray:
        a = np.zeros(3 * self.npoint)
        _L().fdmgpu_cluster_points(self._h, a)
        return a.reshape(self.npoint, 3)

    def profile(self):
        lb1 = np.zeros(self.nligne, np.int32)
        lb2 = np.zeros(self.nligne, np.int32)
        _L().fdmgpu_cluster_profile(self._h, lb1, lb2)
        return lb1, lb2

    def assemble(self, V, Enervide: float, Eimag: float, Ecinetic_out: float, E_comp: bool, basereel: bool):
        """Compressed rows and the incoming-wave RHS of one energy point."""
        cal = bool(E_comp) or not basereel
        n, nl = self.nligne, self.nlmso
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        re = np.zeros(self.nnz)
        im = np.zeros(self.nnz) if cal else None
        rr = np.zeros(n * nl)
        ri = np.zeros(n * nl) if cal else None
        rc = _L().fdmgpu_cluster_assemble(self._h, np.ascontiguousarray(V, np.float64), Enervide, Eimag,
                                          Ecinetic_out, int(bool(E_comp)), int(bool(basereel)), rp, col, re,
                                          _lib.ptr(im), _lib.ptr(rr), _lib.ptr(ri))
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assemble failed ({rc})")
        vals = re + 1j * im if cal else re
        rhs = (rr + 1j * ri).reshape(n, nl) if cal else rr.reshape(n, nl)
        return rp, col, vals, rhs
