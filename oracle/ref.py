"""ORACLE — test infrastructure only.

ctypes harness over ``oracle/_ref/libfdmnes_ref.so``: the reference FDMNES
Fortran compiled from its own sources by ``oracle/Makefile.ref`` (flang).  It
calls the reference's own ``newind`` (mat.f90:1761), ``phiso`` (mat.f90:739),
``cbessneu``/``cbessneur`` (sphere.f90:7226/7356), ``cylm``/``ylmcr``
(general.f90:11606/11692) and ``mat_solve`` (mat_solve_gaussian.f90:5, which
calls the reference ``calcMatRow``, mat.f90:1034) on arrays of a cluster built
by the oracle, so the restatement in ``fdm_oracle.c`` is pinned against the
reference itself.  All Fortran arguments are passed by reference; logicals are
4-byte integers (1 = .true.).

A vanishing pivot makes the reference ``STOP``; callers that expect one run
the call in a child process.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libfdmnes_ref.so")
NOPSM = 64  # main.f90:46

_lib = None


def available() -> bool:
    return os.path.exists(REF_LIB)


def lib():
    global _lib
    if _lib is None:
        # fdmx.f90 calls ZLOG, a non-standard intrinsic flang leaves external;
        # lazy binding leaves that out-of-path symbol unresolved and unused.
        # Global scope: a drop-in mat_solve loaded later resolves the
        # reference's calcMatRow and write_error from here.
        _lib = C.CDLL(REF_LIB, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
    return _lib


class _Args:
    """Keeps every by-reference argument alive for the duration of a call."""

    def __init__(self):
        self.keep = []

    def i(self, v):
        a = np.array([v], np.int32)
        self.keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    def d(self, v):
        a = np.array([v], np.float64)
        self.keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    def b(self, v):
        return self.i(1 if v else 0)

    def arr(self, a, dtype):
        a = np.asfortranarray(np.asarray(a, dtype))
        if a.size == 0:
            a = np.zeros(1, dtype)
        self.keep.append(a)
        return a.ctypes.data_as(C.c_void_p)


def cbessneu(fnorm: complex, z: complex, lmax: int):
    A = _Args()
    b = np.zeros(lmax + 1, np.complex128)
    n = np.zeros(lmax + 1, np.complex128)
    lib().cbessneu_(A.arr([fnorm], np.complex128), A.arr([z], np.complex128), A.i(lmax), A.i(lmax),
                    b.ctypes.data_as(C.c_void_p), n.ctypes.data_as(C.c_void_p))
    return b, n


def cbessneur(fnorm: float, z: float, lmax: int):
    A = _Args()
    b = np.zeros(lmax + 1)
    n = np.zeros(lmax + 1)
    lib().cbessneur_(A.d(fnorm), A.d(z), A.i(lmax), A.i(lmax), b.ctypes.data_as(C.c_void_p),
                     n.ctypes.data_as(C.c_void_p))
    return b, n


def ylm_real(lmax: int, v, r: float):
    A = _Args()
    nlmc = ((lmax + 1) * (lmax + 2)) // 2
    nlmr = (lmax + 1) ** 2
    yc = np.zeros(nlmc, np.complex128)
    yr = np.zeros(nlmr)
    lib().cylm_(A.i(lmax), A.arr(v, np.float64), A.d(r), yc.ctypes.data_as(C.c_void_p), A.i(nlmc))
    lib().ylmcr_(A.i(lmax), A.i(nlmc), A.i(nlmr), yc.ctypes.data_as(C.c_void_p), yr.ctypes.data_as(C.c_void_p))
    return yr


def newind(cl):
    """Reference newind on an oracle cluster (natome = 0, nspino = 1)."""
    A = _Args()
    n = cl.nligne
    lb1 = np.zeros(n, np.int32)
    lb2 = np.zeros(n, np.int32)
    newinv = np.zeros(n, np.int32)
    newp = np.zeros(cl.npoint + 1, np.int32)
    isrt, _ = cl.boundary
    numia = np.zeros(cl.npso, np.int32)
    numia[cl.npoint:] = -2
    xyz = cl.xyz.T  # (4, npso) Fortran order
    lib().newind_(A.arr([0.0], np.float64), A.arr([0], np.int32), A.arr([0], np.int32), A.i(0), A.b(False),
                  A.arr(isrt, np.int32), A.arr(cl.ivois, np.int32), lb1.ctypes.data_as(C.c_void_p),
                  lb2.ctypes.data_as(C.c_void_p), A.i(0), A.i(1), A.i(0), A.arr([0], np.int32), A.i(1),
                  newp.ctypes.data_as(C.c_void_p), newinv.ctypes.data_as(C.c_void_p), A.i(n),
                  A.arr([0], np.int32), A.i(cl.nlmso), A.i(cl.npoint), A.i(cl.nsortf), A.i(1), A.i(cl.npso),
                  A.i(cl.nsort), A.arr(numia, np.int32), A.i(cl.nvois), A.arr(xyz, np.float64))
    return lb1, lb2, newinv, newp


def phiso(cl, E_comp: bool, Ecinetic_out: float, Eimag: float, Enervide: float = 0.0):
    """Reference phiso, non-Rydberg branch.  Returns (Bessel, Neuman) as
    (nsort, lmaxso+1) arrays (complex when E_comp)."""
    A = _Args()
    isrt, _ = cl.boundary
    ns, lm = cl.nsort, cl.lmaxso
    if E_comp:
        bes = np.zeros((ns, lm + 1), np.complex128, order="F")
        neu = np.zeros((ns, lm + 1), np.complex128, order="F")
        besr = np.zeros(1)
        neur = np.zeros(1)
        nsc, nsr = ns, 0
    else:
        bes = np.zeros(1, np.complex128)
        neu = np.zeros(1, np.complex128)
        besr = np.zeros((ns, lm + 1), order="F")
        neur = np.zeros((ns, lm + 1), order="F")
        nsc, nsr = 0, ns
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().phiso_(A.d(0.0), p(bes), p(besr), A.b(E_comp), A.d(Ecinetic_out), A.d(0.0), A.d(Eimag), A.b(False),
                 A.d(Enervide), A.i(0), A.i(1), A.arr(isrt, np.int32), A.i(lm), A.i(0), p(neu), p(neur),
                 A.i(cl.npso), A.i(ns), A.i(nsc), A.i(nsr), A.i(1), A.i(ns), A.d(0.0), A.d(0.0), A.b(False),
                 A.d(0.0), A.arr(cl.xyz.T, np.float64))
    return (bes, neu) if E_comp else (besr, neur)


def mat_solve(cl, V, Enervide, Eimag, bes, neu, E_comp, cal_comp, basereel, rhs=None, fn=None, gradvr=None,
              cgrad=None, ylm_comp=False):
    """Reference mat_solve (93 arguments, mat_solve_gaussian.f90:5-12) with
    the reference calcMatRow assembling each row.  Returns the solution
    (nligne, nlmso), real or complex.  ``fn`` calls another procedure with
    the same argument surface instead (a drop-in replacement).  With gradvr
    (npoint, 3) and cgrad (nvois,) the rows are scalar-relativistic
    (Relativiste = T, nicm = npoint).  ylm_comp: complex harmonics in the
    outer sphere (Ylm_comp = T)."""
    A = _Args()
    n, nl, ns, lm = cl.nligne, cl.nlmso, cl.nsort, cl.lmaxso
    lb1, lb2, newinv, newp = cl.profile()
    isrt, poidso = cl.boundary
    numia = np.zeros(cl.npso, np.int32)
    numia[cl.npoint:] = -2
    nvois = cl.nvois
    smr = np.zeros((nl, n), order="F")
    smi = np.zeros((nl, n), order="F") if cal_comp else np.zeros((1, 1), order="F")
    if rhs is not None:
        smr[:, :] = np.real(rhs).T
        if cal_comp:
            smi[:, :] = np.imag(rhs).T
    if E_comp:
        Bc, Nc = np.asfortranarray(bes, np.complex128), np.asfortranarray(neu, np.complex128)
        Br, Nr = np.zeros(1), np.zeros(1)
        nsc, nsr = ns, 0
    else:
        Bc, Nc = np.zeros(1, np.complex128), np.zeros(1, np.complex128)
        Br, Nr = np.asfortranarray(bes, np.float64), np.asfortranarray(neu, np.float64)
        nsc, nsr = 0, ns
    kar = np.ones((NOPSM, 1), order="F")
    kari = np.zeros((NOPSM, 1), order="F")
    lso = np.array([L for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    mso = np.array([m for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    iso = np.ones(nl, np.int32)
    isvois = np.ones((cl.npso, nvois), np.int32, order="F")
    tf = np.zeros(1)
    tt = np.zeros(1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rel = gradvr is not None
    gv = np.asfortranarray(np.asarray(gradvr, np.float64).reshape(cl.npoint, 3)) if rel else np.zeros(3)
    cg = np.asarray(cgrad, np.float64) if rel else np.zeros(nvois)
    args = [
        A.b(False), A.b(basereel), p(Bc), p(Br), A.b(cal_comp), A.arr(cg, np.float64),
        A.arr(cl.clapl, np.float64), A.b(E_comp), A.d(Eimag), A.d(Enervide), p(gv),
        A.arr([0], np.int32), A.arr([0], np.int32), A.arr([0], np.int32), A.i(0), A.i(1), A.i(1), A.i(0),
        A.b(False), A.arr([0], np.int32), A.arr(iso, np.int32), A.i(1), A.arr(isrt, np.int32),
        A.arr(isvois, np.int32), A.arr(cl.ivois, np.int32), A.arr(kar, np.float64), A.arr(kari, np.float64),
        A.arr([0], np.int32), A.arr(lb1, np.int32), A.arr(lb2, np.int32), A.i(lm), A.arr(lso, np.int32),
        A.arr([0], np.int32), A.i(1), A.i(0), A.arr(mso, np.int32), A.i(1), A.i(1), A.i(0), A.i(1),
        A.arr([0], np.int32), A.arr([0], np.int32), A.i(1), p(Nc), p(Nr), A.arr(newp, np.int32),
        A.arr(newinv, np.int32), A.i(1), A.i(cl.npoint if rel else 1), A.i(cl.npoint), A.i(n),
        A.i(n if cal_comp else 0),
        A.i(cl.nligneso), A.i(1), A.i(1), A.i(1), A.i(nl), A.arr([0], np.int32), A.i(nl),
        A.i(nl if cal_comp else 0), A.i(1), A.i(1), A.i(cl.npoint), A.i(cl.npso), A.i(max(cl.nsortf, 1)),
        A.i(nl), A.i(ns), A.i(nsc), A.i(nsr), A.i(cl.nsortf), A.i(1), A.i(1), A.i(1), A.i(1), A.i(ns),
        A.arr(numia, np.int32), A.i(nvois), A.arr(np.zeros(1), np.float64), A.arr(np.zeros(1), np.float64),
        A.arr(poidso, np.float64), A.b(rel), A.b(False), A.arr(np.ones(cl.npoint), np.float64), p(smi), p(smr),
        A.b(False), A.i(0), p(tf), p(tt), A.arr(np.asarray(V, np.float64), np.float64), A.b(ylm_comp),
        A.arr(np.zeros(1), np.float64), A.arr(cl.ylmso, np.float64),
    ]
    assert len(args) == 93
    (fn or lib().mat_solve_)(*args)
    x = smr.T.copy()
    if cal_comp:
        x = x + 1j * smi.T
    return x


def invcomp(A):
    """Reference ``invcomp`` (mat.f90:1927) with is = 0: the vendored
    ZGETRF/ZGETRI.  A vanishing pivot makes it STOP, so callers pass only
    non-singular matrices here."""
    A_ = _Args()
    a = np.asfortranarray(np.asarray(A, np.complex128)).copy(order="F")
    n = a.shape[0]
    lib().invcomp_(A_.i(n), a.ctypes.data_as(C.c_void_p), A_.i(n), A_.i(n), A_.i(0), A_.b(True))
    return a


def opsym(op: int):
    """The reference's matrix of symmetry operation ``op`` (opsym,
    general.f90:4587): a point p goes to M p."""
    A = _Args()
    m = np.zeros((3, 3), order="F")
    lib().opsym_(A.i(op), m.ctypes.data_as(C.c_void_p))
    return np.array(m)


def grid(rsort: float, adimp: float, iord: int = 4, igrpt: int = 1, iopsymr=None):
    """The reference's grid builders for a cubic grid centred on the
    absorber, with no atoms (natome = 0): reseau (general.f90:9439: points,
    radial order, rvol), laplac (:9732: stencil weights and neighbour table)
    and bordure (:9895: outer-sphere boundary points and weights), whose
    lists come back through recup_bordure (:10140, which reads the scratch
    unit bordure writes).  igrpt / iopsymr select the point group
    (posequiv, :8892); the default is group 1 (P1, the identity only).
    rsort, adimp in bohr.  Returns a dict of the reference's arrays
    (0-based point numbers where an index is stored, -1 for none)."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    eps = 1.0e-4
    rmax = rsort / adimp + np.sqrt(2.0) * iord / 2 + eps
    nx = int(np.ceil(rmax)) + 1
    r = np.arange(-nx, nx + 1)
    g = np.sqrt(r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2)
    npsom = int((g <= rmax).sum()) + 8
    if iopsymr is None:
        iopsymr = np.zeros(NOPSM, np.int32)
        iopsymr[0] = 1
    indice = np.zeros((npsom, 3), np.int32, order="F")
    xyz = np.zeros((4, npsom), order="F")
    rvol = np.zeros(npsom)
    numia = np.zeros(npsom, np.int32)
    mpres = np.zeros((2 * nx + 1,) * 3, np.int32, order="F")
    npoint, npr, npso = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)
    lib().reseau_(A.d(adimp), A.b(False), A.d(0.0), A.b(False), A.i(1), A.i(0), A.i(igrpt), p(indice),
                  A.arr(iopsymr, np.int32), A.i(iord), A.arr([0], np.int32), A.b(False), A.i(0), p(mpres), A.i(0),
                  A.i(1), A.i(npsom), p(npoint), p(npr), p(npso), A.i(npsom), A.i(0), p(numia), A.i(nx),
                  A.arr(np.zeros((3, 1)), np.float64), A.arr(np.zeros((3, 1)), np.float64),
                  A.arr(np.zeros(1), np.float64), A.d(rsort), p(rvol), p(xyz))
    npoint, npso = int(npoint[0]), int(npso[0])
    nvois = 3 * iord
    cgrad = np.zeros(nvois)
    clapl = np.zeros(nvois + 1)
    ivois = np.zeros((npsom, nvois), np.int32, order="F")
    isvois = np.zeros((npsom, nvois), np.int32, order="F")
    lib().laplac_(A.d(adimp), A.b(False), p(cgrad), p(clapl), A.i(0), A.i(igrpt), p(indice),
                  A.arr(iopsymr, np.int32), A.i(iord), p(ivois), p(isvois), A.i(0), p(mpres), A.i(npso), A.i(npsom),
                  A.i(nvois), A.i(nx))
    nbm, nbtm, nsm, nstm = (np.zeros(1, np.int32) for _ in range(4))
    unit = 77
    lib().bordure_(A.b(False), A.i(0), A.arr(iopsymr, np.int32), A.i(iord), A.i(unit), p(ivois), A.i(0), A.i(0),
                   p(nbm), p(nbtm), A.i(npsom), A.i(npoint), A.i(npso), A.i(npsom), p(nsm), p(nstm), p(numia),
                   A.i(nvois), A.arr(np.zeros((3, 1)), np.float64), p(rvol), p(xyz))
    nsm_, nstm_ = max(int(nsm[0]), 1), max(int(nstm[0]), 1)
    isrt = np.zeros(nstm_, np.int32)
    poidso = np.zeros(nsm_)
    nsort, nsortf = np.zeros(1, np.int32), np.zeros(1, np.int32)
    lib().recup_bordure_(A.arr(np.zeros((1, 1), np.int32), np.int32), A.arr(np.zeros((1, 1), np.int32), np.int32),
                         A.i(unit), p(isrt), A.i(1), A.i(0), A.i(0), A.arr([0], np.int32), A.arr([0], np.int32),
                         A.i(max(int(nbm[0]), 1)), A.i(max(int(nbtm[0]), 1)), A.i(nsm_), p(nsort), p(nsortf),
                         A.i(nstm_), A.arr(np.zeros((1, 1)), np.float64), p(poidso))
    nsort, nsortf = int(nsort[0]), int(nsortf[0])
    return dict(npoint=npoint, npso=npso, indice=indice[:npso].copy(), radius=xyz[3, :npso].copy(),
                xyz=xyz[:3, :npso].T.copy(), rvol=rvol[:npoint].copy(), numia=numia[:npso].copy(),
                clapl=clapl, cgrad=cgrad, ivois=ivois[:npso].copy() - 1, isvois=isvois[:npso].copy(),
                nsortf=nsortf, nsort=nsort, isrt=isrt[:nsort].copy() - 1, poidso=poidso[:nsortf].copy())


def atom_system(rsort: float, adimp: float, lmaxso: int, posi, rmt, lmax, iord: int = 4):
    """A cluster with muffin-tin atoms (point group 1, one spin, regular
    solution) built entirely by the reference: reseau with the atoms
    (general.f90:9439, points inside a sphere marked), laplac, bordure and
    recup_bordure (outer and atom border lists, poidso, poidsa), Ylmpt
    (general.f90:11505: Ylmso, Ylmato) and newind (mat.f90:1761: atom blocks
    in the radial order, lb1/lb2, new, newinv).  posi (natome, 3) bohr, the
    absorber at the origin, sorted by distance; rmt (natome,) bohr; lmax
    (natome,).  Returns a dict of the reference's arrays (1-based point and
    row numbers, Fortran shapes)."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    posi = np.asfortranarray(np.asarray(posi, np.float64).reshape(-1, 3).T)  # (3, natome)
    natome = posi.shape[1]
    rmt = np.asarray(rmt, np.float64)
    lmax = np.asarray(lmax, np.int32)
    eps = 1.0e-4
    rmax = rsort / adimp + np.sqrt(2.0) * iord / 2 + eps
    nx = int(np.ceil(rmax)) + 1
    r = np.arange(-nx, nx + 1)
    g = np.sqrt(r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2)
    npsom = int((g <= rmax).sum()) + 8
    iopsymr = np.zeros(NOPSM, np.int32)
    iopsymr[0] = 1
    indice = np.zeros((npsom, 3), np.int32, order="F")
    xyz = np.zeros((4, npsom), order="F")
    rvol = np.zeros(npsom)
    numia = np.zeros(npsom, np.int32)
    mpres = np.zeros((2 * nx + 1,) * 3, np.int32, order="F")
    npoint, npr, npso = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)
    rmt_t = np.concatenate([[0.0], rmt])  # rmt(0:ntype), one type per atom
    lib().reseau_(A.d(adimp), A.b(False), A.d(0.0), A.b(False), A.i(1), A.i(0), A.i(1), p(indice),
                  A.arr(iopsymr, np.int32), A.i(iord), A.arr(np.arange(1, natome + 1), np.int32), A.b(False), A.i(0),
                  p(mpres), A.i(natome), A.i(1), A.i(npsom), p(npoint), p(npr), p(npso), A.i(npsom), A.i(natome),
                  p(numia), A.i(nx), A.arr(np.zeros((3, 1)), np.float64), p(posi), A.arr(rmt_t, np.float64),
                  A.d(rsort), p(rvol), p(xyz))
    npoint, npr, npso = int(npoint[0]), int(npr[0]), int(npso[0])
    nvois = 3 * iord
    cgrad = np.zeros(nvois)
    clapl = np.zeros(nvois + 1)
    ivois = np.zeros((npsom, nvois), np.int32, order="F")
    isvois = np.zeros((npsom, nvois), np.int32, order="F")
    lib().laplac_(A.d(adimp), A.b(False), p(cgrad), p(clapl), A.i(0), A.i(1), p(indice), A.arr(iopsymr, np.int32),
                  A.i(iord), p(ivois), p(isvois), A.i(0), p(mpres), A.i(npso), A.i(npsom), A.i(nvois), A.i(nx))
    nbm, nbtm, nsm, nstm = (np.zeros(1, np.int32) for _ in range(4))
    unit = 78
    lib().bordure_(A.b(False), A.i(0), A.arr(iopsymr, np.int32), A.i(iord), A.i(unit), p(ivois), A.i(0), A.i(natome),
                   p(nbm), p(nbtm), A.i(npsom), A.i(npoint), A.i(npso), A.i(npsom), p(nsm), p(nstm), p(numia),
                   A.i(nvois), p(posi), p(rvol), p(xyz))
    nbm, nbtm, nsm, nstm = (max(int(v[0]), 1) for v in (nbm, nbtm, nsm, nstm))
    ibord = np.zeros((nbtm, natome), np.int32, order="F")
    isbord = np.zeros((nbtm, natome), np.int32, order="F")
    nbord = np.zeros(natome, np.int32)
    nbordf = np.zeros(natome, np.int32)
    poidsa = np.zeros((nbm, natome), order="F")
    isrt = np.zeros(nstm, np.int32)
    poidso = np.zeros(nsm)
    nsort, nsortf = np.zeros(1, np.int32), np.zeros(1, np.int32)
    lib().recup_bordure_(p(ibord), p(isbord), A.i(unit), p(isrt), A.i(1), A.i(0), A.i(natome), p(nbord), p(nbordf),
                         A.i(nbm), A.i(nbtm), A.i(nsm), p(nsort), p(nsortf), A.i(nstm), p(poidsa), p(poidso))
    nsort, nsortf = int(nsort[0]), int(nsortf[0])
    nlmsa = ((lmax + 1) ** 2).astype(np.int32)
    nlmmax = int(nlmsa.max())
    nlmso = (lmaxso + 1) ** 2
    ylmato = np.zeros((nbtm, nlmmax, natome), order="F")
    ylmso = np.zeros((nsort, nlmso), order="F")
    rot = np.zeros((3, 3, natome), order="F")
    for ia in range(natome):
        rot[:, :, ia] = np.eye(3)
    lib().ylmpt_(A.arr(np.arange(1, natome + 1), np.int32), p(ibord), A.i(0), A.arr(iopsymr, np.int32), p(isrt),
                 A.arr(np.concatenate([[lmaxso], lmax]), np.int32), A.i(lmaxso), A.i(natome), A.i(natome), p(nbord),
                 A.i(nbtm), A.i(nlmmax), A.i(nlmso), A.i(nsort), A.i(nstm), A.i(npsom), p(posi), p(rot), p(xyz),
                 p(ylmato), p(ylmso))
    nligne = npr + int(nlmsa.sum()) + nlmso  # mat.f90:150
    lb1 = np.zeros(nligne, np.int32)
    lb2 = np.zeros(nligne, np.int32)
    newinv = np.zeros(nligne, np.int32)
    newp = np.zeros(npoint + 1, np.int32)
    ianew = np.zeros(natome, np.int32)
    distai = np.sqrt((posi ** 2).sum(axis=0))
    lib().newind_(A.arr(distai, np.float64), p(ianew), p(ibord), A.i(0), A.b(False), p(isrt), p(ivois), p(lb1),
                  p(lb2), A.i(0), A.i(1), A.i(natome), p(nbordf), A.i(nbtm), p(newp), p(newinv), A.i(nligne),
                  A.arr(nlmsa, np.int32), A.i(nlmso), A.i(npoint), A.i(nsortf), A.i(1), A.i(npsom), A.i(nstm),
                  p(numia), A.i(nvois), p(xyz))
    return dict(rsort=rsort, adimp=adimp, npoint=npoint, npr=npr, npso=npso, npsom=npsom, natome=natome, posi=posi,
                rmt=rmt, lmax=lmax,
                lmaxso=lmaxso, nlmso=nlmso, nligne=nligne, nligneso=nligne - nlmso, indice=indice, xyz=xyz, rvol=rvol,
                numia=numia, clapl=clapl, cgrad=cgrad, ivois=ivois, isvois=isvois, nsort=nsort, nsortf=nsortf, nsm=nsm,
                nstm=nstm,
                isrt=isrt, poidso=poidso, nbm=nbm, nbtm=nbtm, nbord=nbord, nbordf=nbordf, ibord=ibord, isbord=isbord,
                poidsa=poidsa, nlmsa=nlmsa, nlmmax=nlmmax, ylmato=ylmato, ylmso=ylmso, ianew=ianew, lb1=lb1, lb2=lb2,
                newinv=newinv, new=newp)


def phiso_sys(s, E_comp: bool, Ecinetic_out: float, Eimag: float, Enervide: float = 0.0):
    """Reference phiso (non-Rydberg branch) on an atom_system dict."""
    A = _Args()
    ns, lm = s["nsort"], s["lmaxso"]
    if E_comp:
        bes = np.zeros((ns, lm + 1), np.complex128, order="F")
        neu = np.zeros((ns, lm + 1), np.complex128, order="F")
        besr, neur = np.zeros(1), np.zeros(1)
        nsc, nsr = ns, 0
    else:
        bes, neu = np.zeros(1, np.complex128), np.zeros(1, np.complex128)
        besr = np.zeros((ns, lm + 1), order="F")
        neur = np.zeros((ns, lm + 1), order="F")
        nsc, nsr = 0, ns
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().phiso_(A.d(0.0), p(bes), p(besr), A.b(E_comp), A.d(Ecinetic_out), A.d(0.0), A.d(Eimag), A.b(False),
                 A.d(Enervide), A.i(0), A.i(1), A.arr(s["isrt"], np.int32), A.i(lm), A.i(0), p(neu), p(neur),
                 A.i(s["npsom"]), A.i(ns), A.i(nsc), A.i(nsr), A.i(1), A.i(ns), A.d(0.0), A.d(0.0), A.b(False),
                 A.d(0.0), A.arr(s["xyz"], np.float64))
    return (bes, neu) if E_comp else (besr, neur)


def irregular_layout(s):
    """The reference newind (mat.f90:1761) for the irregular cluster solution
    of an atom_system dict (mat.f90:140-147 with Test_reg = 0: n_ato_RI = 2,
    no outer-sphere rows): nligne, ianew, lb1, lb2, newinv, new."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    natome = s["natome"]
    nligne = s["npr"] + 2 * int(s["nlmsa"].sum())
    lb1 = np.zeros(nligne, np.int32)
    lb2 = np.zeros(nligne, np.int32)
    newinv = np.zeros(nligne, np.int32)
    newp = np.zeros(s["npoint"] + 1, np.int32)
    ianew = np.zeros(natome, np.int32)
    distai = np.sqrt((s["posi"] ** 2).sum(axis=0))
    lib().newind_(A.arr(distai, np.float64), p(ianew), A.arr(s["ibord"], np.int32), A.i(0), A.b(True),
                  A.arr(s["isrt"], np.int32), A.arr(s["ivois"], np.int32), p(lb1), p(lb2), A.i(0), A.i(2),
                  A.i(natome), A.arr(s["nbordf"], np.int32), A.i(s["nbtm"]), p(newp), p(newinv), A.i(nligne),
                  A.arr(s["nlmsa"], np.int32), A.i(s["nlmso"]), A.i(s["npoint"]), A.i(s["nsortf"]), A.i(1),
                  A.i(s["npsom"]), A.i(s["nstm"]), A.arr(s["numia"], np.int32), A.i(s["ivois"].shape[1]),
                  A.arr(s["xyz"], np.float64))
    return dict(nligne=nligne, ianew=ianew, lb1=lb1, lb2=lb2, newinv=newinv, new=newp)


def mat_solve_sys(s, V, Enervide, Eimag, bes, neu, E_comp, cal_comp, basereel, phiato, fn=None, phiato2=None,
                  gradvr=None):
    """Reference mat_solve (with its calcMatRow) on an atom_system dict.
    phiato: list over atoms of (nbord(ia), nlmsa(ia)) arrays, complex
    (nphiato7 = 2) or real.  Returns the solution (nligne, nlmso).  ``fn``
    calls another procedure with the same argument surface instead.  With
    phiato2 (the irregular atomic solutions) it solves the irregular cluster
    solution instead (irregular, n_ato_RI = n_RI = 2, Test_reg = 0, the
    irregular_layout rows).  With gradvr (npoint, 3) the rows are
    scalar-relativistic (Relativiste = T, nicm = npoint, laplac's cgrad)."""
    A = _Args()
    irregular = phiato2 is not None
    rel = gradvr is not None
    gv = np.asfortranarray(np.asarray(gradvr, np.float64).reshape(s["npoint"], 3)) if rel else np.zeros(3)
    if irregular:
        lay = irregular_layout(s)
        s = dict(s, nligne=lay["nligne"], nligneso=lay["nligne"], ianew=lay["ianew"], lb1=lay["lb1"],
                 lb2=lay["lb2"], newinv=lay["newinv"], new=lay["new"])
    n, nl, ns, lm = s["nligne"], s["nlmso"], s["nsort"], s["lmaxso"]
    natome, nbtm, nlmmax = s["natome"], s["nbtm"], s["nlmmax"]
    nvois = s["ivois"].shape[1]
    phi, nphi7 = _phi_array(s, phiato, phiato2)
    nlmsam = nlmmax
    lato = np.zeros((nlmsam, natome, 1), np.int32, order="F")
    mato = np.zeros((nlmsam, natome, 1), np.int32, order="F")
    iato = np.ones((nlmsam, natome, 1), np.int32, order="F")
    for ia in range(natome):
        k = 0
        for L in range(int(s["lmax"][ia]) + 1):
            for m in range(-L, L + 1):
                lato[k, ia, 0], mato[k, ia, 0] = L, m
                k += 1
    smr = np.zeros((nl, n), order="F")
    smi = np.zeros((nl, n), order="F") if cal_comp else np.zeros((1, 1), order="F")
    if E_comp:
        Bc, Nc = np.asfortranarray(bes, np.complex128), np.asfortranarray(neu, np.complex128)
        Br, Nr = np.zeros(1), np.zeros(1)
        nsc, nsr = ns, 0
    else:
        Bc, Nc = np.zeros(1, np.complex128), np.zeros(1, np.complex128)
        Br, Nr = np.asfortranarray(bes, np.float64), np.asfortranarray(neu, np.float64)
        nsc, nsr = 0, ns
    kar = np.ones((NOPSM, 1), order="F")
    kari = np.zeros((NOPSM, 1), order="F")
    lso = np.array([L for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    mso = np.array([m for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    iso = np.ones(nl, np.int32)
    tf, tt = np.zeros(1), np.zeros(1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    npsom = s["npsom"]
    args = [
        A.b(False), A.b(basereel), p(Bc), p(Br), A.b(cal_comp), A.arr(s["cgrad"] if rel else np.zeros(nvois), np.float64),
        A.arr(s["clapl"], np.float64), A.b(E_comp), A.d(Eimag), A.d(Enervide), p(gv),
        A.arr(s["ianew"], np.int32), p(iato), A.arr(s["ibord"], np.int32), A.i(0), A.i(1), A.i(1), A.i(0),
        A.b(irregular), A.arr(s["isbord"], np.int32), A.arr(iso, np.int32), A.i(1), A.arr(s["isrt"], np.int32),
        A.arr(s["isvois"], np.int32), A.arr(s["ivois"], np.int32), A.arr(kar, np.float64), A.arr(kari, np.float64),
        p(lato), A.arr(s["lb1"], np.int32), A.arr(s["lb2"], np.int32), A.i(lm), A.arr(lso, np.int32),
        p(mato), A.i(1), A.i(0), A.arr(mso, np.int32), A.i(2 if irregular else 1), A.i(2 if irregular else 1),
        A.i(natome), A.i(s["nbm"]),
        A.arr(s["nbord"], np.int32), A.arr(s["nbordf"], np.int32), A.i(nbtm), p(Nc), p(Nr),
        A.arr(s["new"], np.int32), A.arr(s["newinv"], np.int32), A.i(1), A.i(s["npoint"] if rel else 1),
        A.i(s["npoint"]), A.i(n),
        A.i(n if cal_comp else 0), A.i(s["nligneso"]), A.i(nlmsam), A.i(nlmmax), A.i(nlmmax), A.i(nl),
        A.arr(s["nlmsa"], np.int32), A.i(nl), A.i(nl if cal_comp else 0), A.i(nbtm), A.i(nphi7),
        A.i(s["npoint"]), A.i(npsom), A.i(max(s["nsm"], 1)), A.i(nl), A.i(ns), A.i(nsc), A.i(nsr),
        A.i(s["nsortf"]), A.i(1), A.i(1), A.i(1), A.i(1), A.i(s["nstm"]), A.arr(s["numia"], np.int32), A.i(nvois),
        p(phi), A.arr(s["poidsa"], np.float64), A.arr(s["poidso"], np.float64), A.b(rel), A.b(False),
        A.arr(s["rvol"][: s["npoint"]], np.float64), p(smi), p(smr), A.b(False), A.i(0), p(tf), p(tt),
        A.arr(np.asarray(V, np.float64), np.float64), A.b(False), A.arr(s["ylmato"], np.float64),
        A.arr(s["ylmso"], np.float64),
    ]
    assert len(args) == 93
    (fn or lib().mat_solve_)(*args)
    x = smr.T.copy()
    if cal_comp:
        x = x + 1j * smi.T
    return x


def spinorbit_layout(s):
    """The reference newind (mat.f90:1761) with nspino = 2 for an
    atom_system dict: two rows per interstitial point, atom blocks of
    2 nlmsa unknowns (L, m, solution index, the index fastest, as lmrep,
    general.f90:11227, lists them in P1) and 2 nlmso outer unknowns (L, m,
    spin): nligne, nlmsa, nlmso, ianew, lb1, lb2, newinv, new."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    natome = s["natome"]
    nlmsa = (2 * s["nlmsa"]).astype(np.int32)
    nlmso = 2 * s["nlmso"]
    nligne = 2 * s["npr"] + int(nlmsa.sum()) + nlmso
    lb1 = np.zeros(nligne, np.int32)
    lb2 = np.zeros(nligne, np.int32)
    newinv = np.zeros(nligne, np.int32)
    newp = np.zeros(s["npoint"] + 1, np.int32)
    ianew = np.zeros(natome, np.int32)
    distai = np.sqrt((s["posi"] ** 2).sum(axis=0))
    lib().newind_(A.arr(distai, np.float64), p(ianew), A.arr(s["ibord"], np.int32), A.i(0), A.b(False),
                  A.arr(s["isrt"], np.int32), A.arr(s["ivois"], np.int32), p(lb1), p(lb2), A.i(0), A.i(1),
                  A.i(natome), A.arr(s["nbordf"], np.int32), A.i(s["nbtm"]), p(newp), p(newinv), A.i(nligne),
                  A.arr(nlmsa, np.int32), A.i(nlmso), A.i(s["npoint"]), A.i(s["nsortf"]), A.i(2),
                  A.i(s["npsom"]), A.i(s["nstm"]), A.arr(s["numia"], np.int32), A.i(s["ivois"].shape[1]),
                  A.arr(s["xyz"], np.float64))
    return dict(nligne=nligne, nlmsa=nlmsa, nlmso=nlmso, ianew=ianew, lb1=lb1, lb2=lb2, newinv=newinv, new=newp)


def mat_solve_so(s, V, Enervide, Eimag, bes, neu, E_comp, basereel, phiato, gradvr, fn=None):
    """Reference mat_solve (with its calcMatRow) on an atom_system dict with
    spin-orbit coupling (Spinorbite = T, nspino = nspinp = 2; one potential,
    nspin = nspinr = 1; complex arithmetic), rows of spinorbit_layout.
    phiato: per atom a (nbord, (lmax+1)^2, 2, 2) array, phiato(ib, (L, m),
    spin component, solution index); gradvr (npoint, 3).  Returns the
    solution (nligne, 2 nlmso), columns (L, m, spin) with the spin fastest.
    ``fn`` calls another procedure with the same argument surface."""
    A = _Args()
    lay = spinorbit_layout(s)
    n, nl, ns, lm = lay["nligne"], lay["nlmso"], s["nsort"], s["lmaxso"]
    natome, nbtm, nlmmax = s["natome"], s["nbtm"], s["nlmmax"]
    nvois = s["ivois"].shape[1]
    nlmsam = 2 * nlmmax
    lato = np.zeros((nlmsam, natome, 1), np.int32, order="F")
    mato = np.zeros((nlmsam, natome, 1), np.int32, order="F")
    iato = np.ones((nlmsam, natome, 1), np.int32, order="F")
    for ia in range(natome):
        k = 0
        for L in range(int(s["lmax"][ia]) + 1):
            for m in range(-L, L + 1):
                for isol in (1, 2):
                    lato[k, ia, 0], mato[k, ia, 0], iato[k, ia, 0] = L, m, isol
                    k += 1
    lso = np.array([L for L in range(lm + 1) for m in range(-L, L + 1) for _ in (1, 2)], np.int32)
    mso = np.array([m for L in range(lm + 1) for m in range(-L, L + 1) for _ in (1, 2)], np.int32)
    iso = np.array([i for L in range(lm + 1) for m in range(-L, L + 1) for i in (1, 2)], np.int32)
    phi = np.zeros((nbtm, nlmmax, 2, 2, 1, natome, 2), order="F")
    for ia, x in enumerate(phiato):
        nb, na = x.shape[:2]
        phi[:nb, :na, :, :, 0, ia, 0] = np.real(x)
        phi[:nb, :na, :, :, 0, ia, 1] = np.imag(x)
    smr = np.zeros((nl, n), order="F")
    smi = np.zeros((nl, n), order="F")
    if E_comp:
        Bc, Nc = np.asfortranarray(bes, np.complex128), np.asfortranarray(neu, np.complex128)
        Br, Nr = np.zeros(1), np.zeros(1)
        nsc, nsr = ns, 0
    else:
        Bc, Nc = np.zeros(1, np.complex128), np.zeros(1, np.complex128)
        Br, Nr = np.asfortranarray(bes, np.float64), np.asfortranarray(neu, np.float64)
        nsc, nsr = 0, ns
    kar = np.ones((NOPSM, 2), order="F")
    kari = np.zeros((NOPSM, 2), order="F")
    gv = np.zeros((s["npoint"], 3, 2), order="F")
    gv[:, :, 0] = gradvr
    gv[:, :, 1] = gradvr
    tf, tt = np.zeros(1), np.zeros(1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    npsom = s["npsom"]
    args = [
        A.b(False), A.b(basereel), p(Bc), p(Br), A.b(True), A.arr(s["cgrad"], np.float64),
        A.arr(s["clapl"], np.float64), A.b(E_comp), A.d(Eimag), A.d(Enervide), p(gv),
        A.arr(lay["ianew"], np.int32), p(iato), A.arr(s["ibord"], np.int32), A.i(0), A.i(1), A.i(1), A.i(0),
        A.b(False), A.arr(s["isbord"], np.int32), A.arr(iso, np.int32), A.i(1), A.arr(s["isrt"], np.int32),
        A.arr(s["isvois"], np.int32), A.arr(s["ivois"], np.int32), A.arr(kar, np.float64), A.arr(kari, np.float64),
        p(lato), A.arr(lay["lb1"], np.int32), A.arr(lay["lb2"], np.int32), A.i(lm), A.arr(lso, np.int32),
        p(mato), A.i(1), A.i(0), A.arr(mso, np.int32), A.i(1), A.i(1), A.i(natome), A.i(s["nbm"]),
        A.arr(s["nbord"], np.int32), A.arr(s["nbordf"], np.int32), A.i(nbtm), p(Nc), p(Nr),
        A.arr(lay["new"], np.int32), A.arr(lay["newinv"], np.int32), A.i(1), A.i(s["npoint"]), A.i(s["npoint"]),
        A.i(n), A.i(n), A.i(n - nl), A.i(nlmsam), A.i(nlmmax), A.i(nlmmax), A.i(s["nlmso"]),
        A.arr(lay["nlmsa"], np.int32), A.i(nl), A.i(nl), A.i(nbtm), A.i(2),
        A.i(s["npoint"]), A.i(npsom), A.i(max(s["nsm"], 1)), A.i(nl), A.i(ns), A.i(nsc), A.i(nsr),
        A.i(s["nsortf"]), A.i(1), A.i(2), A.i(2), A.i(1), A.i(s["nstm"]), A.arr(s["numia"], np.int32), A.i(nvois),
        p(phi), A.arr(s["poidsa"], np.float64), A.arr(s["poidso"], np.float64), A.b(False), A.b(False),
        A.arr(s["rvol"][: s["npoint"]], np.float64), p(smi), p(smr), A.b(True), A.i(0), p(tf), p(tt),
        A.arr(np.asarray(V, np.float64), np.float64), A.b(False), A.arr(s["ylmato"], np.float64),
        A.arr(s["ylmso"], np.float64),
    ]
    assert len(args) == 93
    (fn or lib().mat_solve_)(*args)
    return smr.T + 1j * smi.T


def sphere(s, iapr, r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, V_intmax=1.0e6, n_ri=1):
    """The reference's Sphere (sphere.f90:8) for the atoms of one prototype
    of an atom_system dict (every atom its own prototype: atom iapr, 1-based),
    one spin, spherical potential Vr(r) (Ry) on the radial grid r (bohr),
    atomic number numat, matching radius rmtg, no spin-orbit or Hubbard
    term, the regular solution, Renorm = T, real harmonics.  Returns
    (phiato of the atom, (nbord, (lmax+1)^2) complex, Tau_ato of the atom,
    ((lmax+1)^2,) diagonal); with n_ri = 2 the irregular solution's phiato
    (Sch_radial_solsing, sphere.f90:152-162) comes second."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    natome, nbtm = s["natome"], s["nbtm"]
    nlmagm = (lmax + 1) ** 2
    nlmmax = s["nlmmax"]
    nr = len(r)
    phiato = np.zeros((nbtm, nlmagm, 1, 1, n_ri, natome, 2), order="F")
    tau = np.zeros((nlmagm, 1, nlmagm, 1, natome), np.complex128, order="F")
    iopsymr = np.zeros(NOPSM, np.int32)
    iopsymr[0] = 1
    args = [
        A.arr(np.zeros((3, 1)), np.float64), A.arr([Ecinetic], np.float64), A.d(Eimag), A.d(Enervide), A.d(Enervide),
        A.b(False), A.b(False), A.b(False), A.b(False), A.b(False), A.i(1), A.i(iapr),
        A.arr(np.arange(1, natome + 1), np.int32), A.arr(s["ibord"], np.int32), A.i(0),
        A.arr(np.zeros((natome + 1, 1), np.int32), np.int32), A.arr(np.ones(natome, np.int32), np.int32),
        A.arr(iopsymr, np.int32), A.i(lmax), A.i(0), A.i(0), A.i(1), A.i(natome), A.i(natome), A.i(n_ri), A.i(natome),
        A.arr(s["nbord"], np.int32), A.i(nbtm), A.i(nbtm), A.i(1), A.i(1), A.i(1), A.i(nlmagm), A.i(nlmmax), A.i(nbtm),
        A.i(2), A.i(s["npsom"]), A.i(nr), A.i(1), A.i(1), A.i(1), A.i(numat), p(phiato), A.arr(s["posi"], np.float64),
        A.arr(np.asarray(r, np.float64), np.float64), A.b(relativiste), A.b(True), A.d(rmtg), A.d(rmtg), A.b(False),
        p(tau), A.arr(np.zeros((1, 1, 1, 1), np.complex128), np.complex128), A.d(V_intmax), A.arr([0.0], np.float64),
        A.arr(np.asarray(Vr, np.float64).reshape(nr, 1, 1), np.float64), A.arr(s["xyz"], np.float64), A.b(False),
        A.arr(s["ylmato"], np.float64),
    ]
    assert len(args) == 58
    lib().sphere_(*args)
    ia = iapr - 1
    nb = int(s["nbord"][ia])
    phi = phiato[:nb, :, 0, 0, 0, ia, 0] + 1j * phiato[:nb, :, 0, 0, 0, ia, 1]
    taus = np.array([tau[k, 0, k, 0, ia] for k in range(nlmagm)])
    if n_ri == 2:
        return phi, phiato[:nb, :, 0, 0, 1, ia, 0] + 1j * phiato[:nb, :, 0, 0, 1, ia, 1], taus
    return phi, taus


def _phi_array(s, phiato, phiato2=None):
    """phiato(nphiato1, nlmagm, nspinp, nspino, n_RI, natome, nphiato7) from
    per-atom (nbord, nlmsa) arrays of the regular (and irregular) atomic
    solutions."""
    sets = [phiato] if phiato2 is None else [phiato, phiato2]
    cplx = any(np.iscomplexobj(x) for xs in sets for x in xs)
    nphi7 = 2 if cplx else 1
    phi = np.zeros((s["nbtm"], s["nlmmax"], 1, 1, len(sets), s["natome"], nphi7), order="F")
    for r, xs in enumerate(sets):
        for ia, x in enumerate(xs):
            nb, na = x.shape
            phi[:nb, :na, 0, 0, r, ia, 0] = np.real(x)
            if cplx:
                phi[:nb, :na, 0, 0, r, ia, 1] = np.imag(x)
    return phi, nphi7


MAT_ARGS = (
    "Adimp Ampl_dft Atom_axe Axe_atom_grn Base_hexa Basereelt Cal_xanes cgrad clapl Dist_coop distai E_comp "
    "Ecinetic_out Eclie_out Eimag Eneg Enervide FDM_comp Full_atom gradvr ia_coop iaabsi iaprotoi iato ibord icheck "
    "ie igroupi igrph irep_util isbord iso ispinin isrt ivois isvois karact Lato Lmaxa lmaxso Lso mato mpinodes "
    "mpirank0 mso n_atom_0 n_atom_coop n_atom_ind n_atom_proto n_RI nab_coop natome nbm nbord nbordf nbtm ngroup_m "
    "ngrph nim nicm nlm_rixs nlmagm nlmmax nlmomax nlmsa nlmsam nlmso nphiato1 nphiato7 npoint npr npsom nsm nsort "
    "nsortf nso1 nspin nspino nspinp nstm numia nvois phiato poidsa poidso Posi R_rydb Recop Relativiste Repres_comp "
    "RIXS Rmtg Rsort rvol Rydberg Solsing Spinorbite State_all_r Sym_cubic Tau_ato Tau_coop Tau_RI Taull Test_reg "
    "Time_fill Time_tria V0bd_out Vr xyz Ylm_comp Ylmato Ylmso").split()


def mat_tau(s, adimp, V, E, Eimag, E_comp, fdm_comp, basereel, phiato, phiato2=None, ylm_comp=False):
    """The reference ``mat`` (mat.f90:5) on an atom_system dict: phiso,
    newind and mat_solve (with its calcMatRow) for the regular cluster
    solution and, with FDM_comp (n_RI = 2, Test_reg = 0), the irregular one,
    then the reduction to the atoms' amplitude products: Taull (mat.f90:297-
    405, then the symmetrisation :596-641) or, with complex energy, Tau_RI
    (:229-295).  Point group 1, one spin, every atom (Cal_xanes false).
    phiato / phiato2: per atom (nbord, nlmsa) regular / irregular atomic
    solutions.  Returns (Taull (natome, nlm, nlm), Tau_RI (natome, n_RI,
    nlm, nlm)) with nlm = nlmmax."""
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    natome, nbtm, nlmmax, nl, lm = s["natome"], s["nbtm"], s["nlmmax"], s["nlmso"], s["lmaxso"]
    nvois = s["ivois"].shape[1]
    n_RI = 2 if fdm_comp else 1
    phi, nphi7 = _phi_array(s, phiato, phiato2 if fdm_comp else None)
    nrepm = 12
    lato = np.zeros((nlmmax, natome, 1), np.int32, order="F")
    mato = np.zeros((nlmmax, natome, 1), np.int32, order="F")
    iato = np.ones((nlmmax, natome, 1), np.int32, order="F")
    for ia in range(natome):
        k = 0
        for L in range(int(s["lmax"][ia]) + 1):
            for m in range(-L, L + 1):
                lato[k, ia, 0], mato[k, ia, 0] = L, m
                k += 1
    lso = np.array([L for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    mso = np.array([m for L in range(lm + 1) for m in range(-L, L + 1)], np.int32)
    irep_util = np.zeros((nrepm, 2), np.int32, order="F")
    irep_util[0, :] = 1
    karact = np.zeros((NOPSM, nrepm), np.complex128, order="F")
    karact[:, 0] = 1.0
    taull = np.zeros((nlmmax, 1, nlmmax, 1, natome), np.complex128, order="F")
    tau_ri = np.zeros((nlmmax, 1, nlmmax, 1, n_RI, natome), np.complex128, order="F")
    tau_ato = np.zeros((nlmmax, 1, nlmmax, 1, 1), np.complex128, order="F")
    dz = np.zeros(1, np.complex128)
    tf, tt = np.zeros(1), np.zeros(1)
    distai = np.sqrt((s["posi"] ** 2).sum(axis=0))
    i32, f64 = np.int32, np.float64
    # mat.f90:5-13, argument by argument
    v = dict(
        Adimp=A.d(adimp), Ampl_dft=p(dz), Atom_axe=A.arr(np.zeros(natome, i32), i32),
        Axe_atom_grn=A.arr(np.zeros((3, 1)), f64), Base_hexa=A.b(False), Basereelt=A.b(basereel),
        Cal_xanes=A.b(False), cgrad=A.arr(np.zeros(nvois), f64), clapl=A.arr(s["clapl"], f64),
        Dist_coop=A.arr(np.zeros(2), f64), distai=A.arr(distai, f64), E_comp=A.b(E_comp),
        Ecinetic_out=A.arr([E], f64), Eclie_out=A.d(0.0), Eimag=A.d(Eimag), Eneg=A.b(False), Enervide=A.d(E),
        FDM_comp=A.b(fdm_comp), Full_atom=A.b(False), gradvr=A.arr(np.zeros(3), f64), ia_coop=A.arr([0], i32),
        iaabsi=A.i(1), iaprotoi=A.arr(np.arange(1, natome + 1), i32), iato=p(iato), ibord=A.arr(s["ibord"], i32),
        icheck=A.i(0), ie=A.i(1), igroupi=A.arr(np.arange(1, natome + 1), i32), igrph=A.i(1),
        irep_util=A.arr(irep_util, i32), isbord=A.arr(s["isbord"], i32), iso=A.arr(np.ones(nl, i32), i32),
        ispinin=A.i(1), isrt=A.arr(s["isrt"], i32), ivois=A.arr(s["ivois"], i32), isvois=A.arr(s["isvois"], i32),
        karact=A.arr(karact, np.complex128), Lato=p(lato), Lmaxa=A.arr(s["lmax"], i32), lmaxso=A.i(lm),
        Lso=A.arr(lso, i32), mato=p(mato), mpinodes=A.i(1), mpirank0=A.i(0), mso=A.arr(mso, i32),
        n_atom_0=A.i(1), n_atom_coop=A.i(0), n_atom_ind=A.i(1), n_atom_proto=A.i(natome), n_RI=A.i(n_RI),
        nab_coop=A.i(0), natome=A.i(natome), nbm=A.i(s["nbm"]), nbord=A.arr(s["nbord"], i32),
        nbordf=A.arr(s["nbordf"], i32), nbtm=A.i(nbtm), ngroup_m=A.i(1), ngrph=A.i(1), nim=A.i(s["npoint"]),
        nicm=A.i(1), nlm_rixs=A.i(1), nlmagm=A.i(nlmmax), nlmmax=A.i(nlmmax), nlmomax=A.i(nl),
        nlmsa=A.arr(s["nlmsa"], i32), nlmsam=A.i(nlmmax), nlmso=A.i(nl), nphiato1=A.i(nbtm), nphiato7=A.i(nphi7),
        npoint=A.i(s["npoint"]), npr=A.i(s["npr"]), npsom=A.i(s["npsom"]), nsm=A.i(max(s["nsm"], 1)),
        nsort=A.i(s["nsort"]), nsortf=A.i(s["nsortf"]), nso1=A.i(nl), nspin=A.i(1), nspino=A.i(1), nspinp=A.i(1),
        nstm=A.i(s["nstm"]), numia=A.arr(s["numia"], i32), nvois=A.i(nvois), phiato=p(phi),
        poidsa=A.arr(s["poidsa"], f64), poidso=A.arr(s["poidso"], f64), Posi=A.arr(s["posi"], f64),
        R_rydb=A.d(0.0), Recop=A.b(False), Relativiste=A.b(False), Repres_comp=A.b(False), RIXS=A.b(False),
        Rmtg=A.arr(np.zeros(natome + 1), f64), Rsort=A.d(s["rsort"]), rvol=A.arr(s["rvol"][: s["npoint"]], f64),
        Rydberg=A.b(False), Solsing=A.b(False), Spinorbite=A.b(False), State_all_r=A.b(False),
        Sym_cubic=A.b(False), Tau_ato=p(tau_ato), Tau_coop=p(dz), Tau_RI=p(tau_ri), Taull=p(taull),
        Test_reg=A.i(0), Time_fill=p(tf), Time_tria=p(tt), V0bd_out=A.arr([0.0], f64),
        Vr=A.arr(np.asarray(V, f64), f64), xyz=A.arr(s["xyz"], f64), Ylm_comp=A.b(ylm_comp),
        Ylmato=A.arr(s["ylmato"], f64), Ylmso=A.arr(s["ylmso"], f64))
    args = [v[k] for k in MAT_ARGS]
    lib().mat_(*args)
    return (np.transpose(taull[:, 0, :, 0, :], (2, 0, 1)).copy(),
            np.transpose(tau_ri[:, 0, :, 0, :, :], (3, 2, 0, 1)).copy())


# ---- the absorbing atom's cross-section tensors (tensor.f90) ----

class _Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def f_integr3(r, fct, radius: float) -> float:
    """The reference's f_integr3 (tools.f90:308) with ir0 = 1."""
    A = _Args()
    L = lib()
    L.f_integr3_.restype = C.c_double
    r = np.asarray(r, np.float64)
    return L.f_integr3_(A.arr(r, np.float64), A.arr(fct, np.float64), A.i(1), A.i(len(r)), A.d(radius))


def gaunt_xrc(l, m, lo, mo, li, mi, ylm_comp=False) -> complex:
    """The reference's Gaunt_xrc (tensor.f90:1870)."""
    A = _Args()
    L = lib()
    L.gaunt_xrc_.restype = _Cplx
    v = L.gaunt_xrc_(A.i(l), A.i(m), A.i(lo), A.i(mo), A.i(li), A.i(mi), A.b(ylm_comp))
    return complex(v.re, v.im)


def coef_init(ninit: int, jseuil: int, lseuil: int, nbseuil: int):
    """The reference's coef_init (general.f90:774): coef_g (ninit, 2), m_g
    (ninit, 2), is_g (ninit,)."""
    A = _Args()
    cg = np.zeros((ninit, 2), order="F")
    mg = np.zeros((ninit, 2), np.int32, order="F")
    ig = np.zeros(ninit, np.int32)
    lib().coef_init_(cg.ctypes.data_as(C.c_void_p), ig.ctypes.data_as(C.c_void_p), A.i(ninit), A.i(jseuil),
                     A.i(lseuil), mg.ctypes.data_as(C.c_void_p), A.i(nbseuil))
    return np.ascontiguousarray(cg), np.ascontiguousarray(mg), ig


def tensor_tables(ops=(1,), magnet=False):
    """The reference's tensdd / tensdq / tensqq (general.f90:5899, 6020,
    6127) for the point group of symmetry operations ops (1-based
    numbers): msymdd / msymddi (3, 3), msymdq / msymdqi (3, 3, 3), msymqq /
    msymqqi (3, 3, 3, 3) in C order of the Fortran indices, and ldip / lqua
    set for every element."""
    A = _Args()
    io = np.zeros(NOPSM, np.int32)
    io[np.asarray(ops) - 1] = 1
    out = {}
    for name, shp in (("dd", (3, 3)), ("dq", (3, 3, 3)), ("qq", (3, 3, 3, 3))):
        a = np.zeros(shp, np.int32, order="F")
        b = np.zeros(shp, np.int32, order="F")
        getattr(lib(), f"tens{name}_")(A.arr(io, np.int32), A.b(magnet), a.ctypes.data_as(C.c_void_p),
                                       b.ctypes.data_as(C.c_void_p))
        out[f"msym{name}"] = np.ascontiguousarray(a)
        out[f"msym{name}i"] = np.ascontiguousarray(b)
    out["ldip"] = np.ones(3, np.int32)
    out["lqua"] = np.ones((3, 3), np.int32)
    return out


def tenseur_car(r, Vr, numat, rmtg, rmtsd, lmax, Ecinetic, Eimag, Enervide, Energ, Eseuil, psii, taull, jseuil,
                lseuil, tables, multipoles=(True, True, True), core_resolved=False, relativiste=False, rot_int=None,
                rot_atom_abs=None, V_intmax=1.0e6, ip0=1, ip_max=2):
    """The reference's Tenseur_car (tensor.f90:15) as fdm.f90:3975 calls it
    for the FDM absorption (no TDDFT, optic, NRIXS, Green or FDM_comp
    output, no singular solution, no spin-orbit, one spin, real harmonics):
    Radial_integral of the absorber's radial grid r, potential Vr, core
    functions psii (nbseuil, nr) at edge energies Eseuil, then tens_ab with
    taull (nlm_probe, nlm_probe).  multipoles: (E1E1, E1E2, E2E2).  Returns
    (secdd (ninitr, 3, 3), secdq (ninitr, 3, 3, 3), secqq (ninitr, 3, 3, 3, 3))
    in C order of the Fortran indices."""
    from .oracle import ninit_of
    A = _Args()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    nr = len(r)
    psii = np.atleast_2d(np.asarray(psii, np.float64))
    nbseuil = psii.shape[0]
    ninit, ninit1 = ninit_of(jseuil, lseuil, nbseuil)
    cg, mg, ig = coef_init(ninit, jseuil, lseuil, nbseuil)
    ninitr = ninit if core_resolved else nbseuil
    nlm = (lmax + 1) ** 2
    mult = np.zeros(10, np.int32)
    mult[0], mult[1], mult[5] = [1 if m else 0 for m in multipoles]
    secdd = np.zeros((3, 3, 1, ninitr, 1), np.complex128, order="F")
    secdq = np.zeros((3, 3, 3, ninitr, 1), np.complex128, order="F")
    secqq = np.zeros((3, 3, 3, 3, ninitr, 1), np.complex128, order="F")
    secdo = np.zeros((3, 3, 3, 3, ninitr, 1), np.complex128, order="F")
    sec2 = [np.zeros((3, 3, ninitr, 1), np.complex128, order="F") for _ in range(4)]
    secoo = [np.zeros((3, 1, 3, 1, ninitr, 1), np.complex128, order="F") for _ in range(2)]
    secm = [np.zeros_like(secdd), np.zeros_like(secdo), np.zeros_like(secdq), np.zeros_like(secqq)]
    ta = np.zeros((nlm, nlm, 1, 1, 2, 2, 1), np.complex128, order="F")
    ta[:, :, 0, 0, 0, 0, 0] = taull
    ta[:, :, 0, 0, 1, 1, 0] = taull
    eye = np.eye(3)
    Fm = lambda a: A.arr(np.asarray(a), np.int32)  # C-order (i, j, ...) tables -> Fortran arrays of the same indices
    args = [
        A.arr(cg, np.float64), A.b(core_resolved), A.arr(np.array([[Ecinetic]]), np.float64), A.d(Eimag),
        A.d(Energ), A.arr([Enervide], np.float64), A.arr(np.asarray(Eseuil, np.float64).reshape(nbseuil), np.float64),
        A.b(False), A.b(False), A.b(False), A.b(False), A.b(False), A.b(False), A.b(False), A.b(False),
        A.i(0), A.i(1), A.i(ip_max), A.i(ip0), A.arr(ig, np.int32), A.i(lmax), A.i(0), Fm(tables["ldip"]),
        A.b(False), Fm(np.zeros((3, 3, 3), np.int32)), A.b(False), Fm(tables["lqua"]), A.i(lseuil), A.arr(mg, np.int32),
        A.i(0), A.i(1), A.i(0), A.i(0),
        Fm(tables["msymdd"]), Fm(tables["msymddi"]), Fm(tables["msymdq"]), Fm(tables["msymdqi"]),
        Fm(np.zeros((3, 3, 3, 3), np.int32)), Fm(np.zeros((3, 3, 3, 3), np.int32)),
        Fm(np.zeros((3, 1, 3, 1), np.int32)), Fm(np.zeros((3, 1, 3, 1), np.int32)),
        Fm(tables["msymqq"]), Fm(tables["msymqqi"]), A.arr(mult, np.int32),
        A.i(1), A.i(1), A.i(1), A.i(1), A.i(1), A.i(nbseuil), A.i(1), A.i(1), A.i(1), A.i(ninit1), A.i(ninit),
        A.i(ninitr), A.i(nbseuil), A.i(1), A.i(nlm), A.i(1), A.i(nlm), A.i(nr), A.i(nr), A.i(1), A.i(1), A.i(1),
        A.i(numat), A.arr(psii.T, np.float64), A.arr(np.asarray(r, np.float64), np.float64), A.b(relativiste),
        A.b(True), A.d(rmtg), A.d(rmtsd), A.arr(np.zeros((1, nlm, 1, 1, nbseuil), np.complex128), np.complex128),
        A.arr(eye if rot_atom_abs is None else rot_atom_abs, np.float64),
        A.arr(eye if rot_int is None else rot_int, np.float64),
        p(secdd), p(secm[0]), p(secdo), p(secm[1]), p(secdq), p(secm[2]), p(sec2[0]), p(sec2[1]), p(sec2[2]),
        p(sec2[3]), p(secoo[0]), p(secoo[1]), p(secqq), p(secm[3]),
        A.b(False), A.b(False), A.b(False), A.arr(np.zeros((nlm, 1, nlm, 1, 1), np.complex128), np.complex128), p(ta),
        A.b(False), A.arr(np.zeros((1, 1, 1, 1), np.complex128), np.complex128), A.d(V_intmax),
        A.arr(np.zeros((1, 1)), np.float64), A.arr(np.asarray(Vr, np.float64).reshape(nr, 1, 1, 1), np.float64),
        A.b(False),
    ]
    assert len(args) == 101, len(args)
    lib().tenseur_car_(*args)
    dd = np.stack([secdd[:, :, 0, k, 0] for k in range(ninitr)])
    dq = np.stack([secdq[:, :, :, k, 0] for k in range(ninitr)])
    qq = np.stack([secqq[:, :, :, :, k, 0] for k in range(ninitr)])
    return dd, dq, qq
