"""Per-energy row assembly on the device (``fdmgpu_assembler_*`` in
include/fdmgpu.h): the energy-independent tables of a cluster are copied to
the GPU once, then each call writes the compressed-row values (and the
incoming-wave RHS) of up to 16 energy points straight into device buffers
the batched solver reads.  Stands in for the per-row calcMatRow calls of
mat.f90:1034-1727 (grid rows :1142-1232, outer columns :1312-1398, outer
rows :1401-1511; with muffin-tin atoms also the couplings to the atoms'
expansions :1233-1309 and the atom rows :1515-1632, from the caller's
phiato; with scalar-relativistic rows the terms of :1155-1156, :1171-1185
from the caller's potential gradient)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .batch import Ordered, library_stream

MAX_POINTS = 16
_V = C.c_void_p


class Tables(C.Structure):
    _fields_ = [("npoint", C.c_int), ("nligne", C.c_int), ("nlmso", C.c_int), ("lmaxso", C.c_int),
                ("nsortf", C.c_int), ("nb", C.c_int), ("nrow_out", C.c_int), ("nlink", C.c_int),
                ("nnz", C.c_int64), ("clapl0", C.c_double), ("pw", C.c_double),
                ("row_ptr", _V), ("base", _V), ("qdiag", _V), ("orow", _V), ("oq", _V), ("optr", _V),
                ("ob", _V), ("ow", _V), ("ylm", _V), ("brad", _V), ("lharm", _V), ("wso", _V), ("dself", _V)]


class AtomTables(C.Structure):
    _fields_ = [("natome", C.c_int), ("nar", C.c_int), ("nal", C.c_int), ("nphi", C.c_int64),
                ("aq", _V), ("aa", _V), ("aptr", _V), ("alphi", _V), ("alw", _V), ("ianew", _V), ("nlmsa", _V),
                ("nbordf", _V), ("phoff", _V), ("bq", _V), ("poidsa", _V), ("ylmato", _V), ("nri", C.c_int)]


class RelTables(C.Structure):
    _fields_ = [("npoint", C.c_int), ("nrow_out", C.c_int), ("nlink", C.c_int), ("nar", C.c_int), ("nal", C.c_int),
                ("clapl", C.c_double * 13), ("cgrad", C.c_double * 13),
                ("rq", _V), ("rfac", _V), ("rdiv", _V), ("opoint", _V), ("odir", _V), ("ofac", _V),
                ("apoint", _V), ("adir", _V)]


def _sigs(L):
    if getattr(L, "_asm_sigs", False):
        return
    Pd, Pv = C.POINTER(C.c_double), C.POINTER(_V)
    L.fdmgpu_assembler_run_ri.argtypes = [_V, C.c_int, Pd, Pd, Pd, C.c_int, C.c_int, C.c_int, Pv, Pv, Pv, C.c_int, Pv,
                                          Pv]
    P = C.POINTER
    L.fdmgpu_assembler_create.argtypes = [P(Tables), C.c_int, _V, P(_V)]
    L.fdmgpu_assembler_destroy.argtypes = [_V]
    L.fdmgpu_assembler_destroy.restype = None
    L.fdmgpu_assembler_run.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_int, C.c_int,
                                       C.c_int, P(_V), P(_V), P(_V)]
    L.fdmgpu_assembler_create_atoms.argtypes = [P(Tables), P(AtomTables), C.c_int, _V, P(_V)]
    L.fdmgpu_assembler_run_atoms.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_int,
                                             C.c_int, C.c_int, P(_V), P(_V), C.c_int, P(_V), P(_V)]
    L.fdmgpu_assembler_set_relativistic.argtypes = [_V, P(RelTables)]
    L.fdmgpu_assembler_run_rel.argtypes = [_V, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_int,
                                           C.c_int, C.c_int, P(_V), P(_V), P(_V), C.c_int, P(_V), P(_V)]
    L._asm_sigs = True


def _dptrs(ts):
    arr = (_V * len(ts))()
    for k, t in enumerate(ts):
        if t is None:
            arr[k] = None
            continue
        if not t.is_contiguous():
            raise ValueError("device buffers must be contiguous")
        arr[k] = t.data_ptr()
    return arr


class DeviceAssembler:
    """Assembles the systems of one cluster (``fdmnes_amd.cluster.Cluster``)
    for many energies on the GPU.  ``run`` fills complex128 (or float64 in
    real arithmetic) value tensors of length nnz in the cluster's compressed
    row order, and optionally (nligne, nlmso) RHS tensors.  Runs on a stream
    of its own, ordered after the caller's queued work and before its later
    work (see ``fdmnes_amd.batch.Ordered``)."""

    def __init__(self, cluster, max_points: int = MAX_POINTS, stream: torch.cuda.Stream | None = None,
                 relativistic: bool = False):
        """relativistic: scalar-relativistic rows (the reference's
        Relativiste); ``run`` then takes the potential's gradient G."""
        L = _lib.lib()
        _sigs(L)
        self.L = L
        t = cluster.assembly_tables()
        self.nnz, self.nligne, self.nlmso, self.npoint = t["nnz"], t["nligne"], t["nlmso"], t["npoint"]
        names = ("row_ptr", "base", "qdiag", "orow", "oq", "optr", "ob", "ow", "ylm", "brad", "lharm", "wso", "dself")
        self._keep = {k: np.ascontiguousarray(t[k]) for k in names}
        tb = Tables(t["npoint"], t["nligne"], t["nlmso"], t["lmaxso"], t["nsortf"], t["nb"], t["nrow_out"],
                    t["nlink"], t["nnz"], t["clapl0"], t["pw"], *[self._keep[k].ctypes.data_as(_V) for k in names])
        self.stream = library_stream(stream)
        h = _V()
        self.natome = int(getattr(cluster, "natome", 0))
        self.irregular = False
        if self.natome:
            from .cluster import ATOM_TABLES
            at = cluster.atom_assembly_tables()
            self.nphi = at["nphi"]
            self._keep_at = {k: np.ascontiguousarray(at[k]) for k in ATOM_TABLES}
            self.irregular = bool(getattr(cluster, "irregular", False))
            atb = AtomTables(at["natome"], at["nar"], at["nal"], at["nphi"],
                             *[self._keep_at[k].ctypes.data_as(_V) for k in ATOM_TABLES], 2 if self.irregular else 1)
            _lib.check(L.fdmgpu_assembler_create_atoms(C.byref(tb), C.byref(atb), int(max_points),
                                                       _V(self.stream.cuda_stream), C.byref(h)))
        else:
            _lib.check(L.fdmgpu_assembler_create(C.byref(tb), int(max_points), _V(self.stream.cuda_stream),
                                                 C.byref(h)))
        self._h = h
        self.max_points = int(max_points)
        self.cx = bool(getattr(cluster, "ylm_comp", False))
        if self.cx:  # complex harmonics / characters: the imaginary parts of the tables
            P = C.POINTER(C.c_double)
            L.fdmgpu_assembler_set_complex.argtypes = [_V, P, P, P, P, P]
            self._keep_cx = {k: np.ascontiguousarray(t[k]) for k in ("base_im", "ow_im", "ylm_im", "dself",
                                                                      "dself_im")}
            _lib.check(L.fdmgpu_assembler_set_complex(h, *[self._keep_cx[k].ctypes.data_as(P) for k in
                                                           ("base_im", "ow_im", "ylm_im", "dself", "dself_im")]))
        self.relativistic = bool(relativistic)
        if self.relativistic:
            from .cluster import REL_TABLES
            rt = cluster.rel_tables()
            self._keep_rel = {k: np.ascontiguousarray(rt[k]) for k in REL_TABLES}
            rtb = RelTables(rt["npoint"], rt["nrow_out"], rt["nlink"], rt["nar"], rt["nal"],
                            (C.c_double * 13)(*rt["clapl"]), (C.c_double * 13)(*rt["cgrad"]),
                            *[self._keep_rel[k].ctypes.data_as(_V) for k in REL_TABLES])
            _lib.check(L.fdmgpu_assembler_set_relativistic(h, C.byref(rtb)))

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.fdmgpu_assembler_destroy(self._h)
            self._h = None

    def run(self, V, energies, E_comp: bool, basereel: bool, vals, rhs=None, update_only: bool = False, phi=None,
            G=None, phi2=None):
        """V: list of float64 device tensors (npoint); energies: list of
        (Enervide, Eimag, Ecinetic_out) in Ry; vals (and rhs) lists of
        contiguous device tensors, complex128 when the arithmetic is complex
        (E_comp or not basereel, or complex harmonics), float64 otherwise.  update_only: the
        buffers hold an earlier call's output for this cluster (same
        arithmetic branch); only the energy- and V-dependent entries are
        rewritten.  phi (clusters with atoms): per point a device tensor
        of the nphi phiato values (Cluster.assemble_atoms's blocks,
        concatenated), complex128 (only in complex arithmetic) or float64.
        G (relativistic assemblers only): per point the float64 device
        gradient of V, (npoint, 3)."""
        n = len(energies)
        if n == 0:
            return
        if not (len(V) == len(vals) == n) or (rhs is not None and len(rhs) != n):
            raise ValueError("one V, vals (and rhs) per energy point")
        cal = bool(E_comp) or not basereel or self.cx
        for t in list(vals) + list(rhs or []):
            if t.is_complex() != cal:
                raise ValueError("value and RHS buffers must be complex exactly when the arithmetic is")
        for t in vals:
            if t.numel() != self.nnz:
                raise ValueError(f"value buffers hold nnz = {self.nnz} entries")
        for t in rhs or []:
            if t.numel() != self.nligne * self.nlmso:
                raise ValueError("RHS buffers are (nligne, nlmso)")
        for t in V:
            if t.dtype != torch.float64 or t.numel() != self.npoint:
                raise ValueError("V: float64, npoint entries")
        if (phi is not None) != bool(self.natome):
            raise ValueError("phi is given exactly for a cluster with atoms")
        if (phi2 is not None) != self.irregular:
            raise ValueError("phi2 (the irregular phiato) is given exactly for the irregular cluster solution")
        if phi2 is not None and (len(phi2) != n or any(t.numel() != self.nphi or t.dtype != phi[0].dtype
                                                       for t in phi2)):
            raise ValueError("phi2: one per energy point, like phi")
        if (G is not None) != self.relativistic:
            raise ValueError("G is given exactly for a relativistic assembler")
        if G is not None:
            if len(G) != n:
                raise ValueError("one G per energy point")
            for t in G:
                if t.dtype != torch.float64 or t.numel() != 3 * self.npoint:
                    raise ValueError("G: float64, (npoint, 3)")
        phi_cplx = False
        if phi is not None:
            if len(phi) != n:
                raise ValueError("one phi per energy point")
            phi_cplx = phi[0].is_complex()
            for t in phi:
                if t.is_complex() != phi_cplx or t.numel() != self.nphi or t.dtype not in (torch.float64,
                                                                                          torch.complex128):
                    raise ValueError(f"phi: nphi = {self.nphi} float64 or complex128 entries, one kind for all")
            if phi_cplx and not cal:
                raise ValueError("complex radial solutions need complex arithmetic (E_comp or not basereel)")
        E = np.array(energies, np.float64).reshape(n, 3)
        D = C.POINTER(C.c_double)
        ev = [np.ascontiguousarray(E[:, k]) for k in range(3)]
        with Ordered(self.stream):
            for k0 in range(0, n, self.max_points):
                k1 = min(n, k0 + self.max_points)
                if self.irregular:
                    rc = self.L.fdmgpu_assembler_run_ri(
                        self._h, k1 - k0, *[np.ascontiguousarray(a[k0:k1]).ctypes.data_as(D) for a in ev],
                        int(bool(E_comp)), int(bool(basereel)), int(bool(update_only)), _dptrs(V[k0:k1]),
                        _dptrs(phi[k0:k1]), _dptrs(phi2[k0:k1]), int(phi_cplx), _dptrs(vals[k0:k1]),
                        _dptrs(rhs[k0:k1]) if rhs is not None else None)
                else:
                    rc = self.L.fdmgpu_assembler_run_rel(
                        self._h, k1 - k0, *[np.ascontiguousarray(a[k0:k1]).ctypes.data_as(D) for a in ev],
                        int(bool(E_comp)), int(bool(basereel)), int(bool(update_only)), _dptrs(V[k0:k1]),
                        _dptrs(G[k0:k1]) if G is not None else None,
                        _dptrs(phi[k0:k1]) if phi is not None else None, int(phi_cplx), _dptrs(vals[k0:k1]),
                        _dptrs(rhs[k0:k1]) if rhs is not None else None)
                if rc:
                    break
        _lib.check(rc)
