"""ORACLE — test infrastructure only.

ctypes wrapper over ``oracle/liboracle.so``, the plain-C restatement of the
FDMNES finite-difference path (see ``fdm_oracle.c`` for the file:line map).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only to check results.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_cluster_new.restype = C.c_void_p
        L.orc_cluster_new.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_cluster_new_atoms.restype = C.c_void_p
        L.orc_cluster_new_atoms.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip]
        L.orc_cluster_atom_sizes.argtypes = [C.c_void_p, _ip]
        L.orc_cluster_atom.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_cluster_numia.argtypes = [C.c_void_p, _ip]
        L.orc_cluster_set_phiato.restype = C.c_int
        L.orc_cluster_set_phiato.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p]
        L.orc_cluster_new_atoms_ri.restype = C.c_void_p
        L.orc_cluster_new_atoms_ri.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip,
                                               C.c_int]
        L.orc_cluster_set_phiato_ri.restype = C.c_int
        L.orc_cluster_set_phiato_ri.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_void_p]
        L.orc_cluster_set_relativistic.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_cluster_cgrad.argtypes = [C.c_void_p, _dp]
        L.orc_cluster_set_ylm_comp.argtypes = [C.c_void_p, C.c_int]
        L.orc_sphere.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_void_p,
                                 C.c_void_p]
        L.orc_cluster_free.argtypes = [C.c_void_p]
        L.orc_sphere_ri.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                    _dp, _dp, C.c_void_p, C.c_void_p]
        L.orc_f_integr3.argtypes = [_dp, _dp, C.c_int, C.c_double]
        L.orc_f_integr3.restype = C.c_double
        L.orc_coef_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip]
        L.orc_cgc.argtypes = [C.c_int] * 5
        L.orc_cgc.restype = C.c_double
        L.orc_gaunt_r.argtypes = [C.c_int] * 6
        L.orc_gaunt_r.restype = C.c_double
        L.orc_gaunt_xrc.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_radial_integral.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                          _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_tenseur.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_cluster_sizes.argtypes = [C.c_void_p, _ip]
        L.orc_cluster_xyz.argtypes = [C.c_void_p, _dp]
        L.orc_cluster_indice.argtypes = [C.c_void_p, _ip]
        L.orc_cluster_ivois.argtypes = [C.c_void_p, _ip]
        L.orc_cluster_boundary.argtypes = [C.c_void_p, _ip, _dp]
        L.orc_cluster_ylmso.argtypes = [C.c_void_p, _dp]
        L.orc_cluster_clapl.argtypes = [C.c_void_p, _dp]
        L.orc_cluster_profile.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip]
        L.orc_phiso.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                C.c_int, C.c_int, _dp, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.orc_assemble_dense.restype = C.c_int
        L.orc_assemble_dense.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                         C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_assemble_csr.restype = C.c_int
        L.orc_assemble_csr.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_mat_solve_csr.restype = C.c_int
        L.orc_mat_solve_csr.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _lp, _ip, _dp, C.c_void_p, _dp,
                                        C.c_void_p, C.POINTER(C.c_int)]
        L.orc_dense_solve.restype = C.c_int
        L.orc_dense_solve.argtypes = [C.c_int, C.c_int, _dp, C.c_void_p, _dp, C.c_void_p]
        L.orc_cbessneu.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_cbessneur.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_ylm_real.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.orc_invcomp.restype = C.c_int
        L.orc_invcomp.argtypes = [C.c_int, _dp, _dp]
        L.orc_tau.restype = C.c_int
        L.orc_tau.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_void_p, C.c_int,
                              _ip, _ip, _dp, _dp, C.c_void_p, C.c_void_p]
        L.orc_tau_sym.argtypes = [C.c_int, _ip, C.c_int, _dp, _dp]
        L.orc_tau_ri.argtypes = [C.c_int, _dp, _dp, _ip, _dp, _dp, _ip, C.c_int, _ip, C.c_int, C.c_int, C.c_int,
                                 _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Cluster:
    """Grid, stencil, boundary and profile of one spherical FDM cluster.

    ``rsort`` and ``adimp`` in bohr (reseau, general.f90:9439)."""

    def __init__(self, rsort: float, adimp: float, lmaxso: int, iord: int = 4, atoms=None, irregular=False):
        """atoms: None, or (posi (natome, 3) bohr sorted by distance, rmt
        (natome,) bohr, lmax (natome,)) for muffin-tin atoms (P1).
        irregular: the rows of the irregular cluster solution (atoms only)."""
        L = lib()
        self.irregular = bool(irregular)
        if atoms is None:
            self._h = L.orc_cluster_new(rsort, adimp, iord, lmaxso)
            self.natome = 0
        else:
            posi, rmt, lmax = atoms
            posi = np.ascontiguousarray(posi, np.float64).reshape(-1, 3)
            self.natome = len(posi)
            self._h = L.orc_cluster_new_atoms_ri(rsort, adimp, iord, lmaxso, self.natome, posi.reshape(-1),
                                                 np.ascontiguousarray(rmt, np.float64),
                                                 np.ascontiguousarray(lmax, np.int32), int(self.irregular))
        if not self._h:
            raise MemoryError("orc_cluster_new failed")
        s = np.zeros(7, np.int32)
        L.orc_cluster_sizes(self._h, s)
        (self.npoint, self.npso, self.nligne, self.nlmso, self.nsortf, self.nsort, self.nvois) = map(int, s)
        self.lmaxso = lmaxso
        self.nligneso = self.nligne if self.irregular else self.nligne - self.nlmso

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_cluster_free(self._h)
            self._h = None

    @property
    def numia(self):
        a = np.zeros(self.npso, np.int32)
        lib().orc_cluster_numia(self._h, a)
        return a

    def atom(self, ia: int) -> dict:
        """Atom ia (1-based): nbord, nbordf, nlmsa, ianew, ibord (1-based
        points), poidsa, ylmato (nbord, nlmsa)."""
        L = lib()
        s = np.zeros(4, np.int32)
        L.orc_cluster_atom_sizes(self._h, s)
        nbtm, nlmmax = int(s[2]), int(s[3])
        info = np.zeros(4, np.int32)
        ib = np.zeros(nbtm, np.int32)
        pa = np.zeros(nbtm)
        ya = np.zeros(nbtm * nlmmax)
        L.orc_cluster_atom(self._h, ia, info, ib, pa, ya)
        nb, nbf, nl, ianew = map(int, info)
        return dict(nbord=nb, nbordf=nbf, nlmsa=nl, ianew=ianew, ibord=ib[:nb].copy(), poidsa=pa[:nb].copy(),
                    ylmato=ya[: nb * nl].reshape(nb, nl).copy())

    @property
    def npr(self):
        s = np.zeros(4, np.int32)
        lib().orc_cluster_atom_sizes(self._h, s)
        return int(s[1])

    def set_phiato(self, phi, phi2=None):
        """phi: list over atoms of (nbord, nlmsa) arrays, complex (nphiato7 = 2)
        or real; phi2 the same for the irregular atomic solutions (the
        irregular cluster solution)."""
        sets = [phi] if phi2 is None else [phi, phi2]
        cplx = any(np.iscomplexobj(p) for xs in sets for p in xs)
        for i_ri, xs in enumerate(sets, 1):
            flat = np.concatenate([np.asarray(p).reshape(-1) for p in xs]) if xs else np.zeros(1)
            re = np.ascontiguousarray(np.real(flat), np.float64)
            im = np.ascontiguousarray(np.imag(flat), np.float64) if cplx else None
            if lib().orc_cluster_set_phiato_ri(self._h, i_ri, int(cplx), re, _ptr(im)):
                raise MemoryError("orc_cluster_set_phiato failed")

    def set_relativistic(self, gradvr):
        """Scalar-relativistic rows (Relativiste, mat.f90:1155-1156,
        :1171-1185) with gradvr (npoint, 3), the potential's gradient at the
        interior points (Ry/bohr); None switches them off."""
        if gradvr is None:
            lib().orc_cluster_set_relativistic(self._h, 0, None)
            return
        g = np.asfortranarray(np.asarray(gradvr, np.float64).reshape(self.npoint, 3))
        if lib().orc_cluster_set_relativistic(self._h, 1, g.ctypes.data_as(C.c_void_p)):
            raise MemoryError("orc_cluster_set_relativistic failed")

    def set_ylm_comp(self, on: bool):
        """Complex harmonics in the outer sphere (Ylm_comp); needs complex
        arithmetic (cal_comp) in every solve and assembly."""
        if lib().orc_cluster_set_ylm_comp(self._h, int(bool(on))):
            raise ValueError("Ylm_comp is restated for clusters without atoms")

    @property
    def cgrad(self):
        a = np.zeros(self.nvois)
        lib().orc_cluster_cgrad(self._h, a)
        return a

    @property
    def xyz(self):
        a = np.zeros(4 * self.npso)
        lib().orc_cluster_xyz(self._h, a)
        return a.reshape(self.npso, 4)  # row p = (x, y, z, r) of point p+1

    @property
    def indice(self):
        a = np.zeros(3 * self.npso, np.int32)
        lib().orc_cluster_indice(self._h, a)
        return a.reshape(3, self.npso).T.copy()

    @property
    def ivois(self):
        a = np.zeros(self.npso * self.nvois, np.int32)
        lib().orc_cluster_ivois(self._h, a)
        return a.reshape(self.nvois, self.npso).T.copy()

    @property
    def boundary(self):
        isrt = np.zeros(self.nsort, np.int32)
        poidso = np.zeros(max(self.nsortf, 1))
        lib().orc_cluster_boundary(self._h, isrt, poidso)
        return isrt, poidso[: self.nsortf]

    @property
    def ylmso(self):
        a = np.zeros(self.nsort * self.nlmso)
        lib().orc_cluster_ylmso(self._h, a)
        return a.reshape(self.nlmso, self.nsort).T.copy()

    @property
    def clapl(self):
        a = np.zeros(self.nvois + 1)
        lib().orc_cluster_clapl(self._h, a)
        return a

    def profile(self):
        lb1 = np.zeros(self.nligne, np.int32)
        lb2 = np.zeros(self.nligne, np.int32)
        newinv = np.zeros(self.nligne, np.int32)
        newp = np.zeros(self.npoint + 1, np.int32)
        lib().orc_cluster_profile(self._h, lb1, lb2, newinv, newp)
        return lb1, lb2, newinv, newp

    def phiso(self, E_comp: bool, Ecinetic_out: float, Eimag: float):
        n = self.nsort * (self.lmaxso + 1) * (2 if E_comp else 1)
        b = np.zeros(max(n, 1))
        nn = np.zeros(max(n, 1))
        lib().orc_phiso(self._h, int(E_comp), Ecinetic_out, Eimag, b, nn)
        if E_comp:
            b = b[: 2 * self.nsort * (self.lmaxso + 1)].view(np.complex128)
            nn = nn[: 2 * self.nsort * (self.lmaxso + 1)].view(np.complex128)
        return b.reshape(self.lmaxso + 1, self.nsort).T, nn.reshape(self.lmaxso + 1, self.nsort).T

    def solve(self, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel, rhs=None,
              inc_wave=True, row_stop=1):
        """mat_solve restatement on calcMatRow rows.  rhs: (nligne, nlmso)
        complex or real initial RHS (zero if None).  Returns (status, x,
        fail_row, t_fill, t_tria, nnz)."""
        V = np.ascontiguousarray(V, np.float64)
        n, nl = self.nligne, self.nlmso
        smr = np.zeros(n * nl)
        smi = np.zeros(n * nl) if cal_comp else None
        if rhs is not None:
            smr[:] = np.real(rhs).reshape(-1)
            if cal_comp:
                smi[:] = np.imag(rhs).reshape(-1)
        fr, tf, tt, nnz = C.c_int(0), C.c_double(0), C.c_double(0), C.c_int64(0)
        st = lib().orc_solve(self._h, V, Enervide, Eimag, Ecinetic_out, int(E_comp), int(cal_comp),
                             int(basereel), int(inc_wave), smr, _ptr(smi), int(row_stop), C.byref(fr),
                             C.byref(tf), C.byref(tt), C.byref(nnz))
        x = smr.reshape(n, nl) if not cal_comp else (smr + 1j * smi).reshape(n, nl)
        return st, x, fr.value, tf.value, tt.value, nnz.value

    def assemble_dense(self, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel):
        n, nl = self.nligne, self.nlmso
        ar, ai = np.zeros(n * n), np.zeros(n * n)
        rr, ri = np.zeros(n * nl), np.zeros(n * nl)
        st = lib().orc_assemble_dense(self._h, np.ascontiguousarray(V, np.float64), Enervide, Eimag,
                                      Ecinetic_out, int(E_comp), int(cal_comp), int(basereel), ar, ai, rr, ri)
        if st:
            raise ValueError(f"orc_assemble_dense: {st}")
        A = (ar + 1j * ai).reshape(n, n).T  # column-major -> A[row, col]
        b = (rr + 1j * ri).reshape(n, nl)
        return (A, b) if cal_comp else (A.real.copy(), b.real.copy())

    def assemble_csr(self, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel):
        L = lib()
        V = np.ascontiguousarray(V, np.float64)
        nnz = C.c_int64(0)
        L.orc_assemble_csr(self._h, V, Enervide, Eimag, Ecinetic_out, int(E_comp), int(cal_comp),
                           int(basereel), None, None, None, None, None, None, C.byref(nnz))
        n, nl = self.nligne, self.nlmso
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        vr, vi = np.zeros(nnz.value), np.zeros(nnz.value)
        rr, ri = np.zeros(n * nl), np.zeros(n * nl)
        st = L.orc_assemble_csr(self._h, V, Enervide, Eimag, Ecinetic_out, int(E_comp), int(cal_comp),
                                int(basereel), _ptr(rp), _ptr(col), _ptr(vr), _ptr(vi), _ptr(rr), _ptr(ri),
                                C.byref(nnz))
        if st:
            raise ValueError(f"orc_assemble_csr: {st}")
        vals = vr + 1j * vi if cal_comp else vr
        b = (rr + 1j * ri).reshape(n, nl) if cal_comp else rr.reshape(n, nl)
        return rp, col, vals, b


def mat_solve_csr(lb1, lb2, row_ptr, col, vals, rhs, cal_comp):
    """Generic mat_solve restatement (1-based lb1/lb2/col).  rhs (n, nrhs)."""
    n, nl = rhs.shape
    smr = np.ascontiguousarray(np.real(rhs), np.float64).reshape(-1).copy()
    smi = np.ascontiguousarray(np.imag(rhs), np.float64).reshape(-1).copy() if cal_comp else None
    vr = np.ascontiguousarray(np.real(vals), np.float64)
    vi = np.ascontiguousarray(np.imag(vals), np.float64) if cal_comp else None
    fr = C.c_int(0)
    st = lib().orc_mat_solve_csr(n, nl, int(cal_comp), np.ascontiguousarray(lb1, np.int32),
                                 np.ascontiguousarray(lb2, np.int32), np.ascontiguousarray(row_ptr, np.int64),
                                 np.ascontiguousarray(col, np.int32), vr, _ptr(vi), smr, _ptr(smi), C.byref(fr))
    x = (smr + 1j * smi).reshape(n, nl) if cal_comp else smr.reshape(n, nl)
    return st, x, fr.value


def dense_solve(A, B):
    """Partial-pivoting dense LU (brute force).  A (n,n), B (n,nrhs)."""
    n = A.shape[0]
    nrhs = B.shape[1]
    ar = np.ascontiguousarray(np.real(A).T).reshape(-1).copy()
    ai = np.ascontiguousarray(np.imag(A).T).reshape(-1).copy()
    br = np.ascontiguousarray(np.real(B).T).reshape(-1).copy()
    bi = np.ascontiguousarray(np.imag(B).T).reshape(-1).copy()
    st = lib().orc_dense_solve(n, nrhs, ar, _ptr(ai), br, _ptr(bi))
    if st:
        raise np.linalg.LinAlgError("singular")
    return (br + 1j * bi).reshape(nrhs, n).T


def cbessneu(fnorm: complex, z: complex, lmax: int):
    b = np.zeros(2 * (lmax + 1))
    n = np.zeros(2 * (lmax + 1))
    lib().orc_cbessneu(fnorm.real, fnorm.imag, z.real, z.imag, lmax, b, n)
    return b.view(np.complex128), n.view(np.complex128)


def cbessneur(fnorm: float, z: float, lmax: int):
    b = np.zeros(lmax + 1)
    n = np.zeros(lmax + 1)
    lib().orc_cbessneur(fnorm, z, lmax, b, n)
    return b, n


def ylm_real(lmax: int, v, r: float):
    y = np.zeros((lmax + 1) ** 2)
    lib().orc_ylm_real(lmax, np.ascontiguousarray(v, np.float64), r, y)
    return y


def invcomp(A):
    """invcomp restatement (mat.f90:1927, ZGETRF + ZGETRI).  Returns
    (inverse, info); info = 1-based first zero pivot, 0 on success."""
    a = np.asfortranarray(np.asarray(A, np.complex128))
    n = a.shape[0]
    re = np.ascontiguousarray(a.real.T).reshape(-1).copy()  # column-major
    im = np.ascontiguousarray(a.imag.T).reshape(-1).copy()
    info = lib().orc_invcomp(n, re, im)
    return (re + 1j * im).reshape(n, n).T.copy(), info


def tau(x, nligneso: int, basereel: bool, fdm_comp: bool, atoms, tau0=None, want_norm=False):
    """Post-solve reduction of one point (mat.f90:297-405).  x (nligne,
    nlmso) real or complex solution; atoms = [(first_row, nrows), ...]
    (0-based).  Returns (list of per-atom (na, na) complex tau, norm or None,
    info); tau0 (same structure) is accumulated into like Taull."""
    x = np.asarray(x)
    cal = np.iscomplexobj(x)
    n, nl = x.shape
    xr = np.ascontiguousarray(x.real, np.float64).reshape(-1)
    xi = np.ascontiguousarray(x.imag, np.float64).reshape(-1) if cal else None
    ianew = np.array([a[0] for a in atoms], np.int32)
    nlmsa = np.array([a[1] for a in atoms], np.int32)
    tot = int(sum(int(k) * int(k) for k in nlmsa))
    tr, ti = np.zeros(max(tot, 1)), np.zeros(max(tot, 1))
    if tau0 is not None:
        flat = np.concatenate([np.asarray(t, np.complex128).reshape(-1) for t in tau0]) if tot else np.zeros(0)
        tr[:tot], ti[:tot] = flat.real, flat.imag
    nr = np.zeros(nl * nl) if want_norm else None
    ni = np.zeros(nl * nl) if want_norm else None
    info = lib().orc_tau(n, nl, int(nligneso), int(cal), int(bool(basereel)), int(bool(fdm_comp)), xr, _ptr(xi),
                         len(atoms), ianew, nlmsa, tr, ti, _ptr(nr), _ptr(ni))
    out, off = [], 0
    for k in nlmsa:
        k = int(k)
        out.append((tr[off:off + k * k] + 1j * ti[off:off + k * k]).reshape(k, k))
        off += k * k
    norm = (nr + 1j * ni).reshape(nl, nl).T.copy() if want_norm else None
    return out, norm, info


def sphere(r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, bxyz, pos, ylmato, V_intmax=1.0e6):
    """orc_sphere: phiato (nbord, nlm) complex at the atom's border points
    bxyz (nbord, 3) with their real harmonics ylmato (nbord, nlm), and the
    atomic amplitudes Tau of L = 0..lmax; the flag of complex radial
    solutions."""
    ylm = np.ascontiguousarray(ylmato, np.float64)
    nb, nlm = ylm.shape
    pr, pi = np.zeros(nb * nlm), np.zeros(nb * nlm)
    tr, ti = np.zeros(lmax + 1), np.zeros(lmax + 1)
    rc = lib().orc_sphere(len(r), np.ascontiguousarray(r, np.float64), np.ascontiguousarray(Vr, np.float64),
                          int(numat), float(rmtg), int(lmax), int(bool(relativiste)), float(Ecinetic), float(Eimag),
                          float(Enervide), float(V_intmax), nb, np.ascontiguousarray(bxyz, np.float64).reshape(-1),
                          np.ascontiguousarray(pos, np.float64), ylm.reshape(-1), nlm, pr, pi, _ptr(tr), _ptr(ti))
    if rc < 0:
        raise ValueError("orc_sphere: bad argument")
    return (pr + 1j * pi).reshape(nb, nlm), tr + 1j * ti, bool(rc)


def sphere_ri(r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, bxyz, pos, ylmato, V_intmax=1.0e6):
    """orc_sphere_ri: as sphere, with the irregular solution's phiato:
    (phi, phi2, tau, complex flag)."""
    ylm = np.ascontiguousarray(ylmato, np.float64)
    nb, nlm = ylm.shape
    pr, pi, qr, qi = (np.zeros(nb * nlm) for _ in range(4))
    tr, ti = np.zeros(lmax + 1), np.zeros(lmax + 1)
    rc = lib().orc_sphere_ri(len(r), np.ascontiguousarray(r, np.float64), np.ascontiguousarray(Vr, np.float64),
                             int(numat), float(rmtg), int(lmax), int(bool(relativiste)), float(Ecinetic), float(Eimag),
                             float(Enervide), float(V_intmax), nb, np.ascontiguousarray(bxyz, np.float64).reshape(-1),
                             np.ascontiguousarray(pos, np.float64), ylm.reshape(-1), nlm, pr, pi, qr, qi, _ptr(tr),
                             _ptr(ti))
    if rc < 0:
        raise ValueError("orc_sphere_ri: bad argument")
    return (pr + 1j * pi).reshape(nb, nlm), (qr + 1j * qi).reshape(nb, nlm), tr + 1j * ti, bool(rc)


def tau_sym(taus, ylm_comp: bool = False):
    """The symmetrisation of Taull at the end of mat (mat.f90:596-641, one
    spin): taus a list of per-atom (na, na) complex arrays (harmonics in
    (L, m) order); returns new arrays."""
    nlmsa = np.array([t.shape[0] for t in taus], np.int32)
    tr = np.ascontiguousarray(np.concatenate([np.asarray(t).real.reshape(-1) for t in taus]), np.float64)
    ti = np.ascontiguousarray(np.concatenate([np.asarray(t).imag.reshape(-1) for t in taus]), np.float64)
    lib().orc_tau_sym(len(taus), nlmsa, int(bool(ylm_comp)), tr, ti)
    out, off = [], 0
    for k in nlmsa:
        k = int(k)
        out.append((tr[off:off + k * k] + 1j * ti[off:off + k * k]).reshape(k, k))
        off += k * k
    return out


def tau_ri(x_reg, x_irr, ianew_reg, ianew_irr, nlmsa, n_ato_ri: int = 2, n_ri: int = 2, round_regular=True,
           tau0=None):
    """Tau_RI of mat.f90:229-295 from the regular and irregular cluster
    solutions (complex (nligne, nlmso) each; ianew 0-based first rows of the
    atom blocks).  Returns per atom (n_ri, na, na) complex arrays.  With
    round_regular (the reference's arithmetic) the regular solution enters
    rounded to single precision, as mat.f90:211 stores it.  tau0 (the same
    structure) is added to, as Tau_RI accumulates."""
    xr = np.asarray(x_reg, np.complex128)
    xi = np.asarray(x_irr, np.complex128)
    nl = xr.shape[1]
    nlmsa = np.ascontiguousarray(nlmsa, np.int32)
    tot = int(sum(n_ri * int(k) * int(k) for k in nlmsa))
    o_r, o_i = np.zeros(tot), np.zeros(tot)
    if tau0 is not None:
        flat = np.concatenate([np.asarray(t, np.complex128).reshape(-1) for t in tau0])
        o_r, o_i = np.ascontiguousarray(flat.real), np.ascontiguousarray(flat.imag)
    lib().orc_tau_ri(nl, np.ascontiguousarray(xr.real).reshape(-1), np.ascontiguousarray(xr.imag).reshape(-1),
                     np.ascontiguousarray(ianew_reg, np.int32), np.ascontiguousarray(xi.real).reshape(-1),
                     np.ascontiguousarray(xi.imag).reshape(-1), np.ascontiguousarray(ianew_irr, np.int32),
                     len(nlmsa), nlmsa, int(n_ato_ri), int(n_ri), int(bool(round_regular)), o_r, o_i)
    out, off = [], 0
    for k in nlmsa:
        k = int(k)
        m = n_ri * k * k
        out.append((o_r[off:off + m] + 1j * o_i[off:off + m]).reshape(n_ri, k, k))
        off += m
    return out


# ---- the absorbing atom's cross-section tensors (tensor.f90) ----

def f_integr3(r, fct, radius: float) -> float:
    """orc_f_integr3: the reference's integral of fct from 0 to radius."""
    r = np.ascontiguousarray(r, np.float64)
    return lib().orc_f_integr3(r, np.ascontiguousarray(fct, np.float64), len(r), float(radius))


def ninit_of(jseuil: int, lseuil: int, nbseuil: int):
    """(ninit, ninit1): the number of core states |j, mj> of the edge(s) and
    of the first edge (coef_init's loop, general.f90:798-806)."""
    is_init = 1 if jseuil in (1, 3, 5, 7) else -1
    counts = []
    for k in range(nbseuil):
        if k == 1:
            is_init = -is_init
        counts.append(int(round(2 * (lseuil + 0.5 * is_init) + 1)))
    return sum(counts), counts[0]


def coef_init(jseuil: int, lseuil: int, nbseuil: int):
    """orc_coef_init: coef_g (ninit, 2), m_g (ninit, 2), is_g (ninit,)."""
    ninit, _ = ninit_of(jseuil, lseuil, nbseuil)
    cg, mg, ig = np.zeros(2 * ninit), np.zeros(2 * ninit, np.int32), np.zeros(ninit, np.int32)
    n = lib().orc_coef_init(jseuil, lseuil, nbseuil, ninit, cg, mg, ig)
    if n != ninit:
        raise ValueError("orc_coef_init: inconsistent ninit")
    return cg.reshape(ninit, 2), mg.reshape(ninit, 2), ig


def cgc(l1, l2, l3, m1, m2) -> float:
    return lib().orc_cgc(l1, l2, l3, m1, m2)


def gaunt_r(l1, m1, l2, m2, l3, m3) -> float:
    return lib().orc_gaunt_r(l1, m1, l2, m2, l3, m3)


def gaunt_xrc(l, m, lo, mo, li, mi) -> complex:
    gr, gi = C.c_double(), C.c_double()
    lib().orc_gaunt_xrc(l, m, lo, mo, li, mi, C.byref(gr), C.byref(gi))
    return complex(gr.value, gi.value)


def radial_integral(r, Vr, numat, rmtg, rmtsd, lmax, relativiste, Ecinetic, Eimag, Enervide, Energ, Eseuil, psii,
                    ip0=1, ip_max=2, V_intmax=1.0e6):
    """orc_radial_integral: rof (nbseuil, ip_max - ip0 + 1, lmax + 1) complex
    of the golden rule for core functions psii (nbseuil, nr) at edge
    energies Eseuil; and the flag of complex radial solutions."""
    r = np.ascontiguousarray(r, np.float64)
    psii = np.ascontiguousarray(np.atleast_2d(psii), np.float64)
    nb = psii.shape[0]
    nip = ip_max - ip0 + 1
    re, im = np.zeros(nb * nip * (lmax + 1)), np.zeros(nb * nip * (lmax + 1))
    rc = lib().orc_radial_integral(len(r), r, np.ascontiguousarray(Vr, np.float64), int(numat), float(rmtg),
                                   float(rmtsd), int(lmax), int(bool(relativiste)), float(Ecinetic), float(Eimag),
                                   float(Enervide), float(Energ), float(V_intmax), nb,
                                   np.ascontiguousarray(Eseuil, np.float64).reshape(nb), psii.reshape(-1), ip0, ip_max,
                                   re, im)
    if rc < 0:
        raise ValueError("orc_radial_integral: bad argument")
    return (re + 1j * im).reshape(nb, nip, lmax + 1), bool(rc)


class _TensIn(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("lseuil", "ninit", "ninit1", "nbseuil", "core_resolved", "lmax", "nlm_probe",
                                        "lplus1", "lmoins1", "numat", "e1e1", "e1e2", "e2e2", "ip0", "ip_max")] + \
               [("coef_g", C.c_void_p)] + [(n, C.c_void_p) for n in ("m_g", "ldip", "lqua", "msymdd", "msymddi",
                                                                    "msymdq", "msymdqi", "msymqq", "msymqqi")] + \
               [("rot_int", C.c_void_p), ("rot_atom_abs", C.c_void_p)]


def tenseur(rof, taull, lseuil, coef_g, m_g, ninit1, nbseuil, lmax, tables, multipoles=(True, True, True),
            core_resolved=False, lplus1=False, lmoins1=False, numat=26, rot_int=None, rot_atom_abs=None, ip0=1,
            ip_max=2):
    """orc_tenseur: (secdd (ninitr, 3, 3), secdq (ninitr, 3, 3, 3), secqq
    (ninitr, 3, 3, 3, 3)) complex.  rof (nbseuil, ip_max - ip0 + 1,
    nlm_probe) complex per (L, m) (Radial_integral's rof, the same for every m
    of L); taull (nlm_probe, nlm_probe) complex; tables: dict of ldip (3,),
    lqua (3, 3), msymdd / msymddi (3, 3), msymdq / msymdqi (3, 3, 3), msymqq /
    msymqqi (3, 3, 3, 3) in the reference's index order."""
    rof = np.ascontiguousarray(rof, np.complex128)
    taull = np.ascontiguousarray(taull, np.complex128)
    nlm = taull.shape[0]
    ninit = len(coef_g)
    ninitr = ninit if core_resolved else nbseuil
    keep = {}

    def arr(name, a, dt):
        keep[name] = np.ascontiguousarray(a, dt)
        return keep[name].ctypes.data_as(C.c_void_p)
    eye = np.eye(3)
    d = _TensIn(int(lseuil), ninit, int(ninit1), int(nbseuil), int(bool(core_resolved)), int(lmax), nlm,
                int(bool(lplus1)), int(bool(lmoins1)), int(numat), *[int(bool(m)) for m in multipoles], ip0, ip_max,
                arr("cg", coef_g, np.float64), arr("mg", m_g, np.int32), arr("ldip", tables["ldip"], np.int32),
                arr("lqua", tables["lqua"], np.int32), *[arr(n, tables[n], np.int32) for n in
                                                         ("msymdd", "msymddi", "msymdq", "msymdqi", "msymqq",
                                                          "msymqqi")],
                arr("ri", eye if rot_int is None else rot_int, np.float64),
                arr("ra", eye if rot_atom_abs is None else rot_atom_abs, np.float64))
    dd, dq, qq = np.zeros(ninitr * 18), np.zeros(ninitr * 54), np.zeros(ninitr * 162)
    rc = lib().orc_tenseur(C.byref(d), np.ascontiguousarray(rof.real), np.ascontiguousarray(rof.imag),
                           np.ascontiguousarray(taull.real), np.ascontiguousarray(taull.imag), dd, dq, qq)
    if rc < 0:
        raise ValueError("orc_tenseur: unsupported case")
    c = lambda a, shp: (a[0::2] + 1j * a[1::2]).reshape(shp)
    return c(dd, (ninitr, 3, 3)), c(dq, (ninitr, 3, 3, 3)), c(qq, (ninitr, 3, 3, 3, 3))
