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
        L.fdmgpu_cluster_create_mmm.restype = C.c_int
        L.fdmgpu_cluster_create_mmm.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_create_group_ylm.restype = C.c_int
        L.fdmgpu_cluster_create_group_ylm.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                                       C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_create_group.restype = C.c_int
        L.fdmgpu_cluster_create_group.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                                  C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_rvol.argtypes = [C.c_void_p, _lib.dp]
        L.fdmgpu_cluster_assembly_weights.restype = C.c_int
        L.fdmgpu_cluster_assembly_weights.argtypes = [C.c_void_p, _lib.ip, _lib.dp, _lib.dp]
        L.fdmgpu_cluster_harmonics.argtypes = [C.c_void_p, _lib.ip]
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
        L.fdmgpu_cluster_assemble_rel.restype = C.c_int
        L.fdmgpu_cluster_assemble_rel.argtypes = [C.c_void_p, _lib.dp, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                                  C.c_int, C.c_int, _lib.lp, _lib.ip, _lib.dp, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]
        L.fdmgpu_cluster_assemble_atoms_rel.restype = C.c_int
        L.fdmgpu_cluster_assemble_atoms_rel.argtypes = [C.c_void_p, _lib.dp, C.c_void_p, C.c_double, C.c_double,
                                                        C.c_double, C.c_int, C.c_int, _lib.dp, C.c_void_p,
                                                        C.c_void_p, C.c_void_p, _lib.lp, _lib.ip, _lib.dp,
                                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdmgpu_cluster_assembly_sizes.restype = C.c_int
        L.fdmgpu_cluster_assembly_sizes.argtypes = [C.c_void_p, _lib.lp]
        L.fdmgpu_cluster_assembly_tables_complex.restype = C.c_int
        L.fdmgpu_cluster_assembly_tables_complex.argtypes = [C.c_void_p, _lib.dp, _lib.dp, _lib.lp, _lib.ip, _lib.lp,
                                                             _lib.ip, _lib.ip, _lib.dp, _lib.dp, _lib.dp, _lib.dp,
                                                             _lib.dp, _lib.dp, _lib.dp, _lib.dp]
        L.fdmgpu_cluster_assembly_tables.restype = C.c_int
        L.fdmgpu_cluster_assembly_tables.argtypes = [C.c_void_p, _lib.dp, _lib.lp, _lib.ip, _lib.lp, _lib.ip,
                                                     _lib.ip, _lib.dp, _lib.dp, _lib.dp, _lib.dp]
        L.fdmgpu_cluster_create_atoms.restype = C.c_int
        L.fdmgpu_cluster_create_atoms.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _lib.dp, _lib.dp,
                                                  _lib.ip, C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_create_atoms_ri.restype = C.c_int
        L.fdmgpu_cluster_create_atoms_ri.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _lib.dp, _lib.dp,
                                                     _lib.ip, C.c_int, C.POINTER(C.c_void_p)]
        L.fdmgpu_cluster_assemble_atoms_ri.restype = C.c_int
        L.fdmgpu_cluster_assemble_atoms_ri.argtypes = [C.c_void_p, _lib.dp, C.c_double, C.c_double, C.c_double,
                                                       C.c_int, C.c_int, _lib.dp, C.c_void_p, C.c_void_p,
                                                       C.c_void_p, _lib.lp, _lib.ip, _lib.dp, C.c_void_p,
                                                       C.c_void_p, C.c_void_p]
        L.fdmgpu_cluster_atom_sizes.restype = C.c_int
        L.fdmgpu_cluster_atom_sizes.argtypes = [C.c_void_p, _lib.lp]
        L.fdmgpu_cluster_atom.restype = C.c_int
        L.fdmgpu_cluster_atom.argtypes = [C.c_void_p, C.c_int, _lib.ip, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdmgpu_cluster_point_rows.restype = C.c_int
        L.fdmgpu_cluster_point_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdmgpu_cluster_assemble_atoms.restype = C.c_int
        L.fdmgpu_cluster_assemble_atoms.argtypes = [C.c_void_p, _lib.dp, C.c_double, C.c_double, C.c_double,
                                                    C.c_int, C.c_int, _lib.dp, C.c_void_p, _lib.lp, _lib.ip,
                                                    _lib.dp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdmgpu_cluster_rel_tables.restype = C.c_int
        L.fdmgpu_cluster_rel_tables.argtypes = [C.c_void_p, _lib.lp, _lib.dp, _lib.dp, _lib.ip, _lib.ip, _lib.dp,
                                                _lib.ip, _lib.ip, _lib.dp, _lib.dp]
        L.fdmgpu_cluster_atom_assembly_sizes.restype = C.c_int
        L.fdmgpu_cluster_atom_assembly_sizes.argtypes = [C.c_void_p, _lib.lp]
        L.fdmgpu_cluster_atom_assembly_tables.restype = C.c_int
        L.fdmgpu_cluster_atom_assembly_tables.argtypes = [C.c_void_p, _lib.lp, _lib.ip, _lib.ip, _lib.lp, _lib.dp,
                                                          _lib.ip, _lib.ip, _lib.ip, _lib.lp, _lib.lp, _lib.dp,
                                                          _lib.dp]
        L._cluster_sigs = True
    return L


class Cluster:
    """A spherical cluster of radius ``rsort`` and grid step ``adimp``
    (bohr) with ``(lmaxso+1)**2`` outer-sphere unknowns; with ``irrep`` the
    system of one representation of a subgroup of mmm on its calculation
    area (see fdmgpu_cluster_create_group): ``group`` is the reference's
    point-group number, 8 = mmm (irrep 0..7, wedge x, y, z >= 0, the
    default), 6 = mm2 (irrep 0..3, wedge x, y >= 0), 3 = m (irrep 0 or 4,
    wedge z >= 0), 2 = -1 (irrep 0, 1), 4 = 2 about z (irrep 0, 1),
    5 = 2/m (irrep 0, 1, 4, 5), 7 = 222 (irrep 0..3); irrep is a character
    vector: the operation negating the coordinates in a set of axes has the
    character (-1)^(number of those axes whose bit is set), and irrep is the
    smallest such mask of the representation.  With ``atoms``
    = (posi (natome, 3) bohr sorted by distance from the absorber, rmt
    (natome,) bohr, lmax (natome,)) a P1 cluster with muffin-tin atoms
    (fdmgpu_cluster_create_atoms), assembled with :meth:`assemble_atoms`."""

    # point-group number -> the representation masks it takes
    GROUPS = {1: (), 3: (0, 4), 6: (0, 1, 2, 3), 8: tuple(range(8)), 2: (0, 1), 4: (0, 1), 5: (0, 1, 4, 5),
              7: (0, 1, 2, 3), 9: (0, 1, 2, 3), 10: (0, 1, 2, 3), 11: tuple(range(8))}
    # the four-fold groups 4, -4, 4/m: complex characters and harmonics
    COMPLEX_GROUPS = (9, 10, 11)

    def __init__(self, rsort: float, adimp: float, lmaxso: int, irrep: int = -1, group: int = 8, atoms=None,
                 irregular: bool = False, ylm_comp: bool = False):
        """irregular (with atoms): the rows of the irregular cluster solution
        (fdmgpu_cluster_create_atoms_ri), assembled with
        :meth:`assemble_atoms` given both atomic solutions.  ylm_comp:
        complex harmonics in the outer sphere (the reference's Ylm_comp;
        groups 1-5, always for 4, -4, 4/m: fdmgpu_cluster_create_group_ylm)."""
        L = _L()
        h = C.c_void_p()
        self.irregular = bool(irregular)
        if self.irregular and atoms is None:
            raise ValueError("the irregular cluster solution needs atoms")
        if irrep < 0:
            group = 1
        self.natome = 0
        if atoms is not None:
            if irrep >= 0:
                raise ValueError("clusters with atoms are built in point group P1 only")
            posi, rmt, lmax = atoms
            posi = np.ascontiguousarray(posi, np.float64).reshape(-1, 3)
            self.natome = len(posi)
            rc = L.fdmgpu_cluster_create_atoms_ri(rsort, adimp, lmaxso, self.natome, posi.reshape(-1),
                                                  np.ascontiguousarray(rmt, np.float64),
                                                  np.ascontiguousarray(lmax, np.int32), int(self.irregular),
                                                  C.byref(h))
        else:
            if group in self.COMPLEX_GROUPS and irrep >= 0:
                ylm_comp = True
            rc = L.fdmgpu_cluster_create_group_ylm(rsort, adimp, lmaxso, int(group), int(irrep), int(bool(ylm_comp)),
                                                   C.byref(h))
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_create failed ({rc})")
        self._h = h
        s = np.zeros(6, np.int32)
        L.fdmgpu_cluster_sizes(h, s)
        self.npoint, self.npso, self.nligne, self.nlmso, self.nsortf, self.nsort = map(int, s)
        self.lmaxso = lmaxso
        self.irrep = int(irrep)
        self.group = int(group)
        self.ylm_comp = bool(ylm_comp) and atoms is None
        self.rsort, self.adimp = rsort, adimp
        self.nnz = int(L.fdmgpu_cluster_nnz(h))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _L().fdmgpu_cluster_destroy(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
                pass
            self._h = None

    @property
    def points(self) -> np.ndarray:
        a = np.zeros(3 * self.npoint)
        _L().fdmgpu_cluster_points(self._h, a)
        return a.reshape(self.npoint, 3)

    @property
    def rvol(self) -> np.ndarray:
        a = np.zeros(self.npoint)
        _L().fdmgpu_cluster_rvol(self._h, a)
        return a

    @property
    def harmonics(self) -> np.ndarray:
        a = np.zeros(self.nlmso, np.int32)
        _L().fdmgpu_cluster_harmonics(self._h, a)
        return a

    def profile(self):
        lb1 = np.zeros(self.nligne, np.int32)
        lb2 = np.zeros(self.nligne, np.int32)
        _L().fdmgpu_cluster_profile(self._h, lb1, lb2)
        return lb1, lb2

    def _gradvr(self, gradvr):
        if gradvr is None:
            return None
        g = np.ascontiguousarray(gradvr, np.float64)
        if g.shape != (self.npoint, 3):
            raise ValueError(f"gradvr must be (npoint, 3) = ({self.npoint}, 3), got {g.shape}")
        return g

    def assemble(self, V, Enervide: float, Eimag: float, Ecinetic_out: float, E_comp: bool, basereel: bool,
                 gradvr=None):
        """Compressed rows and the incoming-wave RHS of one energy point.
        gradvr (npoint, 3): the potential's gradient (Ry/bohr) for
        scalar-relativistic rows (the reference's Relativiste).  Complex
        harmonics or characters make the arithmetic complex."""
        cal = bool(E_comp) or not basereel or self.ylm_comp
        n, nl = self.nligne, self.nlmso
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        re = np.zeros(self.nnz)
        im = np.zeros(self.nnz) if cal else None
        rr = np.zeros(n * nl)
        ri = np.zeros(n * nl) if cal else None
        g = self._gradvr(gradvr)
        rc = _L().fdmgpu_cluster_assemble_rel(self._h, np.ascontiguousarray(V, np.float64), _lib.ptr(g), Enervide,
                                              Eimag, Ecinetic_out, int(bool(E_comp)), int(bool(basereel)), rp, col,
                                              re, _lib.ptr(im), _lib.ptr(rr), _lib.ptr(ri))
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assemble failed ({rc})")
        vals = re + 1j * im if cal else re
        rhs = (rr + 1j * ri).reshape(n, nl) if cal else rr.reshape(n, nl)
        return rp, col, vals, rhs

    @property
    def npr(self) -> int:
        """Interstitial grid points (rows of the grid)."""
        s = np.zeros(4, np.int64)
        _L().fdmgpu_cluster_atom_sizes(self._h, s)
        return int(s[1]) if self.natome else self.npoint

    @property
    def numia(self) -> np.ndarray:
        a = np.zeros(self.npoint, np.int32)
        _L().fdmgpu_cluster_point_rows(self._h, _lib.ptr(a), None)
        return a

    @property
    def point_rows(self) -> np.ndarray:
        """0-based row of each interior point, -1 inside an atom."""
        a = np.zeros(self.npoint, np.int32)
        _L().fdmgpu_cluster_point_rows(self._h, None, _lib.ptr(a))
        return a

    def atom(self, a: int) -> dict:
        """Atom a (0-based): nbord, nbordf, nlmsa, ianew (first row, 0-based),
        ibord (0-based points), poidsa (nbordf,), ylmato (nbord, nlmsa)."""
        L = _L()
        info = np.zeros(4, np.int32)
        if L.fdmgpu_cluster_atom(self._h, int(a), info, None, None, None):
            raise IndexError(a)
        nb, nbf, nl, ianew = map(int, info)
        ib = np.zeros(nb, np.int32)
        pa = np.zeros(nbf)
        ya = np.zeros(nb * nl)
        L.fdmgpu_cluster_atom(self._h, int(a), info, _lib.ptr(ib), _lib.ptr(pa), _lib.ptr(ya))
        return dict(nbord=nb, nbordf=nbf, nlmsa=nl, ianew=ianew, ibord=ib, poidsa=pa, ylmato=ya.reshape(nb, nl))

    def assemble_atoms(self, V, Enervide: float, Eimag: float, Ecinetic_out: float, E_comp: bool, basereel: bool,
                       phiato, phiato2=None, gradvr=None):
        """Compressed rows and incoming-wave RHS of one energy point of a
        cluster with atoms; phiato: per atom an (nbord, nlmsa) array of the
        radial solutions times Ylm at its border points, complex or real;
        phiato2 the irregular atomic solutions (irregular clusters only);
        gradvr (npoint, 3) for scalar-relativistic rows."""
        cal = bool(E_comp) or not basereel
        sets = [phiato] if phiato2 is None else [phiato, phiato2]
        cplx = any(np.iscomplexobj(p) for xs in sets for p in xs)
        if cplx and not cal:
            raise ValueError("complex radial solutions need complex arithmetic (E_comp or not basereel)")
        flats = [np.concatenate([np.asarray(p).reshape(-1) for p in xs]) for xs in sets]
        pr = np.ascontiguousarray(flats[0].real, np.float64)
        pi = np.ascontiguousarray(flats[0].imag, np.float64) if cplx else None
        p2r = np.ascontiguousarray(flats[1].real, np.float64) if phiato2 is not None else None
        p2i = np.ascontiguousarray(flats[1].imag, np.float64) if (phiato2 is not None and cplx) else None
        n, nl = self.nligne, self.nlmso
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        re = np.zeros(self.nnz)
        im = np.zeros(self.nnz) if cal else None
        rr = np.zeros(n * nl)
        ri = np.zeros(n * nl) if cal else None
        g = self._gradvr(gradvr)
        rc = _L().fdmgpu_cluster_assemble_atoms_rel(self._h, np.ascontiguousarray(V, np.float64), _lib.ptr(g),
                                                    Enervide, Eimag, Ecinetic_out, int(bool(E_comp)),
                                                    int(bool(basereel)), pr, _lib.ptr(pi), _lib.ptr(p2r),
                                                    _lib.ptr(p2i), rp, col, re, _lib.ptr(im), _lib.ptr(rr),
                                                    _lib.ptr(ri))
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assemble_atoms failed ({rc})")
        vals = re + 1j * im if cal else re
        rhs = (rr + 1j * ri).reshape(n, nl) if cal else rr.reshape(n, nl)
        return rp, col, vals, rhs

    def assembly_tables(self) -> dict:
        """Energy-independent tables of the per-energy assembly
        (fdmgpu_cluster_assembly_tables): numpy arrays plus the sizes."""
        L = _L()
        sz = np.zeros(4, np.int64)
        L.fdmgpu_cluster_assembly_sizes(self._h, sz)
        nrow, nlink, nb, nnz = map(int, sz)
        t = dict(nrow_out=nrow, nlink=nlink, nb=nb, nnz=nnz, npoint=self.npoint, nligne=self.nligne,
                 nlmso=self.nlmso, lmaxso=self.lmaxso, nsortf=self.nsortf,
                 base=np.zeros(nnz), qdiag=np.zeros(self.npoint, np.int64), orow=np.zeros(max(nrow, 1), np.int32),
                 oq=np.zeros(max(nrow, 1), np.int64), optr=np.zeros(nrow + 1, np.int32),
                 ob=np.zeros(max(nlink, 1), np.int32), ow=np.zeros(max(nlink, 1)),
                 ylm=np.zeros(nb * self.nlmso), brad=np.zeros(nb), par=np.zeros(2))
        if self.ylm_comp:  # complex harmonics (and, for 4, -4, 4/m, complex characters)
            t.update(base_im=np.zeros(nnz), ow_im=np.zeros(max(nlink, 1)), ylm_im=np.zeros(nb * self.nlmso),
                     dself_c=np.zeros(3 * self.npoint), dself_im=np.zeros(3 * self.npoint))
            rc = L.fdmgpu_cluster_assembly_tables_complex(self._h, t["base"], t["base_im"], t["qdiag"], t["orow"],
                                                          t["oq"], t["optr"], t["ob"], t["ow"], t["ow_im"], t["ylm"],
                                                          t["ylm_im"], t["brad"], t["par"], t["dself_c"],
                                                          t["dself_im"])
        else:
            rc = L.fdmgpu_cluster_assembly_tables(self._h, t["base"], t["qdiag"], t["orow"], t["oq"], t["optr"],
                                                  t["ob"], t["ow"], t["ylm"], t["brad"], t["par"])
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assembly_tables failed ({rc})")
        t["lharm"] = np.zeros(self.nlmso, np.int32)
        t["wso"] = np.zeros(self.nsortf)
        t["dself"] = np.zeros(3 * self.npoint)
        rc = L.fdmgpu_cluster_assembly_weights(self._h, t["lharm"], t["wso"], t["dself"])
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assembly_weights failed ({rc})")
        if self.ylm_comp:
            t["dself"] = t.pop("dself_c")
        if self.natome:
            phi = [np.zeros((self.atom(a)["nbord"], self.atom(a)["nlmsa"])) for a in range(self.natome)]
            if self.irregular:  # its rows exist with complex energy, in the complex basis, with both solutions
                rp, _, _, _ = self.assemble_atoms(np.zeros(self.npoint), 0.0, 0.01, 1.0, True, False, phi,
                                                  [np.zeros_like(p) for p in phi])
            else:
                rp, _, _, _ = self.assemble_atoms(np.zeros(self.npoint), 0.0, 0.0, 1.0, False, True, phi)
        else:
            rp, _, _, _ = self.assemble(np.zeros(self.npoint), 0.0, 0.0, 1.0, False, True)
        t["row_ptr"] = rp
        t["clapl0"], t["pw"] = float(t["par"][0]), float(t["par"][1])
        return t

    def outer_weights(self) -> np.ndarray:
        """Quadrature weights of the inner outer-sphere points, 4 pi rvol /
        sum(rvol) (bordure)."""
        lh, w, ds = np.zeros(self.nlmso, np.int32), np.zeros(self.nsortf), np.zeros(3 * self.npoint)
        rc = _L().fdmgpu_cluster_assembly_weights(self._h, lh, w, ds)
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_assembly_weights failed ({rc})")
        return w

    def atom_assembly_tables(self) -> dict:
        """The device assembler's atom tables (fdmgpu_cluster_atom_assembly_tables)."""
        if not self.natome:
            raise ValueError("the cluster has no atoms")
        L = _L()
        sz = np.zeros(7, np.int64)
        L.fdmgpu_cluster_atom_assembly_sizes(self._h, sz)
        na, nar, nal, nphi, nrow, nbf, nyl = map(int, sz)
        t = dict(natome=na, nar=nar, nal=nal, nphi=nphi, aq=np.zeros(max(nar, 1), np.int64),
                 aa=np.zeros(max(nar, 1), np.int32), aptr=np.zeros(nar + 1, np.int32),
                 alphi=np.zeros(max(nal, 1), np.int64), alw=np.zeros(max(nal, 1)), ianew=np.zeros(na, np.int32),
                 nlmsa=np.zeros(na, np.int32), nbordf=np.zeros(na, np.int32), phoff=np.zeros(na, np.int64),
                 bq=np.zeros(nrow, np.int64), poidsa=np.zeros(nbf), ylmato=np.zeros(nyl))
        rc = L.fdmgpu_cluster_atom_assembly_tables(self._h, *[t[k] for k in ATOM_TABLES])
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_atom_assembly_tables failed ({rc})")
        return t


    def rel_tables(self) -> dict:
        """The device assembler's tables of scalar-relativistic rows
        (fdmgpu_cluster_rel_tables)."""
        L = _L()
        sz = np.zeros(4, np.int64)
        L.fdmgpu_cluster_assembly_sizes(self._h, sz)
        nrow, nlink = int(sz[0]), int(sz[1])
        nar = nal = 0
        if self.natome:
            sa = np.zeros(7, np.int64)
            L.fdmgpu_cluster_atom_assembly_sizes(self._h, sa)
            nar, nal = int(sa[1]), int(sa[2])
        n12 = 12 * self.npoint
        t = dict(npoint=self.npoint, nrow_out=nrow, nlink=nlink, nar=nar, nal=nal, rq=np.zeros(n12, np.int64),
                 rfac=np.zeros(n12), rdiv=np.zeros(n12), opoint=np.zeros(max(nrow, 1), np.int32),
                 odir=np.zeros(max(nlink, 1), np.int32), ofac=np.zeros(max(nlink, 1)),
                 apoint=np.zeros(max(nar, 1), np.int32), adir=np.zeros(max(nal, 1), np.int32),
                 clapl=np.zeros(13), cgrad=np.zeros(13))
        rc = L.fdmgpu_cluster_rel_tables(self._h, *[t[k] for k in REL_TABLES], t["clapl"], t["cgrad"])
        if rc:
            raise _lib.FdmGpuError(f"fdmgpu_cluster_rel_tables failed ({rc})")
        return t


ATOM_TABLES = ("aq", "aa", "aptr", "alphi", "alw", "ianew", "nlmsa", "nbordf", "phoff", "bq", "poidsa", "ylmato")
REL_TABLES = ("rq", "rfac", "rdiv", "opoint", "odir", "ofac", "apoint", "adir")
