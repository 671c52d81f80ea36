"""Post-solve reduction to the atom amplitude products tau on the device
(``fdmgpu_tau_reduce`` in include/fdmgpu.h): the loop of mat.f90:297-405
after mat_solve (real-basis renormalisation with invcomp, mat.f90:312-343 and
:1927-1987; contraction to Taull, :345-405) for solutions that are already on
the GPU, so only tau (nlmsa^2 complex entries per atom and energy point)
leaves it."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .batch import Ordered, as_real, library_stream

_V = C.c_void_p
_stream = None


class TauDesc(C.Structure):
    _fields_ = [("nligne", C.c_int), ("nlmso", C.c_int), ("nligneso", C.c_int), ("cal_comp", C.c_int),
                ("basereel", C.c_int), ("fdm_comp", C.c_int), ("natome", C.c_int),
                ("ianew", C.POINTER(C.c_int)), ("nlmsa", C.POINTER(C.c_int))]


class TauRIDesc(C.Structure):
    _fields_ = [("nlmso", C.c_int), ("natome", C.c_int), ("ianew_reg", C.POINTER(C.c_int)),
                ("ianew_irr", C.POINTER(C.c_int)), ("nlmsa", C.POINTER(C.c_int)), ("nligne_reg", C.c_int),
                ("nligne_irr", C.c_int), ("n_ato_ri", C.c_int), ("n_ri", C.c_int)]


def _sigs(L):
    if getattr(L, "_tau_sigs", False):
        return
    P = C.POINTER
    L.fdmgpu_tau_reduce.argtypes = [P(TauDesc), C.c_int, P(_V), P(_V), P(_V), C.c_int, _V, _V]
    L.fdmgpu_tau_ri.argtypes = [P(TauRIDesc), C.c_int, P(_V), P(_V), P(_V), C.c_int, _V]
    L.fdmgpu_tau_symmetrize.argtypes = [C.c_int, P(C.c_int), C.c_int, C.c_int, P(_V), _V]
    L._tau_sigs = True


def _lib_stream(stream):
    global _stream
    if stream is not None:
        return library_stream(stream)
    if _stream is None:  # the module's own stream for calls that name none
        _stream = library_stream(None)
    return _stream


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


def tau_reduce(xs, nligneso: int, basereel: bool, fdm_comp: bool, atoms, tau=None, accumulate: bool = False,
               want_norm: bool = False, check: bool = True, stream: torch.cuda.Stream | None = None):
    """Reduces the solutions ``xs`` (list of (nligne, nlmso) device tensors,
    complex128 or float64) of energy points to tau.

    atoms: [(first_row, nrows), ...] 0-based row blocks (ianew, nlmsa).
    tau: None (new zero buffers) or a list of complex128 device tensors of
    sum(nrows^2) entries, added to when ``accumulate``.  Returns (tau,
    norm, info): per point the flat tau tensor (atoms concatenated, each
    block row-major [Lm1][Lm2]), the inverted normalisation (nlmso, nlmso)
    when ``want_norm`` and basereel, and the int32 status tensor.  With
    ``check`` a zero pivot of the normalisation raises FdmGpuError (the
    reference stops there)."""
    L = _lib.lib()
    _sigs(L)
    if not xs:
        return [], [], None
    x0 = xs[0]
    nligne, nlmso = x0.shape
    cal = x0.is_complex()
    for x in xs:
        if x.shape != x0.shape or x.is_complex() != cal or x.dtype not in (torch.complex128, torch.float64):
            raise ValueError("solutions: (nligne, nlmso) complex128 or float64, all alike")
    tot = sum(int(na) * int(na) for _, na in atoms)
    dev = x0.device
    if tau is None:
        tau = [torch.zeros(max(tot, 1), dtype=torch.complex128, device=dev) for _ in xs]
    elif len(tau) != len(xs) or any(t.dtype != torch.complex128 or t.numel() < tot for t in tau):
        raise ValueError(f"tau buffers: complex128, {tot} entries each")
    norm = [torch.empty((nlmso, nlmso), dtype=torch.complex128, device=dev) for _ in xs] \
        if (want_norm and basereel) else None
    info = torch.zeros(len(xs), dtype=torch.int32, device=dev)
    ianew = (C.c_int * max(len(atoms), 1))(*[int(a) for a, _ in atoms])
    nlmsa = (C.c_int * max(len(atoms), 1))(*[int(k) for _, k in atoms])
    d = TauDesc(int(nligne), int(nlmso), int(nligneso), int(cal), int(bool(basereel)), int(bool(fdm_comp)),
                len(atoms), ianew, nlmsa)
    s = _lib_stream(stream)
    with Ordered(s):
        _lib.check(L.fdmgpu_tau_reduce(C.byref(d), len(xs), _dptrs([as_real(x) if cal else x for x in xs]),
                                       _dptrs([as_real(t) for t in tau]),
                                       _dptrs([as_real(t) for t in norm]) if norm else None,
                                       int(bool(accumulate)), _V(info.data_ptr()), _V(s.cuda_stream)))
    if check:
        bad = torch.nonzero(info).flatten().tolist()
        if bad:
            raise _lib.FdmGpuError(f"invcomp: zero pivot {int(info[bad[0]])} in the normalisation of point {bad[0]}")
    return tau, norm, info


def split(t: torch.Tensor, atoms):
    """Views of one point's flat tau as per-atom (nrows, nrows) blocks."""
    out, off = [], 0
    for _, na in atoms:
        out.append(t[off:off + na * na].view(na, na))
        off += na * na
    return out


def tau_ri(x_reg, x_irr, atoms_reg, atoms_irr, tau=None, accumulate: bool = False, n_ato_ri: int = 2,
           n_ri: int = 2, stream: torch.cuda.Stream | None = None):
    """Tau_RI (mat.f90:229-295, FDM_comp) of energy points from their regular
    and irregular cluster solutions (lists of complex128 (nligne, nlmso)
    device tensors; atoms_reg / atoms_irr: [(first_row, nlmsa), ...] of each
    atom in the two systems).  Returns per point a flat complex128 tensor,
    atoms concatenated, each block [i_RI][Lm1][Lm2] (n_ri x nlmsa x nlmsa),
    added to ``tau`` when ``accumulate``."""
    L = _lib.lib()
    _sigs(L)
    if not x_reg:
        return []
    if len(x_irr) != len(x_reg) or len(atoms_irr) != len(atoms_reg):
        raise ValueError("one irregular solution per regular one, the same atoms")
    if any(a[1] != b[1] for a, b in zip(atoms_reg, atoms_irr)):
        raise ValueError("the atoms' nlmsa differ between the two systems")
    for x in list(x_reg) + list(x_irr):
        if x.dtype != torch.complex128 or x.dim() != 2 or x.shape[1] != x_reg[0].shape[1]:
            raise ValueError("solutions: complex128 (nligne, nlmso), one nlmso")
    tot = sum(n_ri * int(na) * int(na) for _, na in atoms_reg)
    dev = x_reg[0].device
    if tau is None:
        tau = [torch.zeros(max(tot, 1), dtype=torch.complex128, device=dev) for _ in x_reg]
    elif len(tau) != len(x_reg) or any(t.dtype != torch.complex128 or t.numel() < tot for t in tau):
        raise ValueError(f"Tau_RI buffers: complex128, {tot} entries each")
    na = max(len(atoms_reg), 1)
    d = TauRIDesc(int(x_reg[0].shape[1]), len(atoms_reg), (C.c_int * na)(*[int(a) for a, _ in atoms_reg]),
                  (C.c_int * na)(*[int(a) for a, _ in atoms_irr]), (C.c_int * na)(*[int(k) for _, k in atoms_reg]),
                  int(x_reg[0].shape[0]), int(x_irr[0].shape[0]), int(n_ato_ri), int(n_ri))
    s = _lib_stream(stream)
    with Ordered(s):
        _lib.check(L.fdmgpu_tau_ri(C.byref(d), len(x_reg), _dptrs([as_real(x) for x in x_reg]),
                                   _dptrs([as_real(x) for x in x_irr]), _dptrs([as_real(t) for t in tau]),
                                   int(bool(accumulate)), _V(s.cuda_stream)))
    return tau


def symmetrize(tau, nlmsa, ylm_comp: bool = False, stream: torch.cuda.Stream | None = None):
    """The symmetrisation ending mat (mat.f90:596-641) in place on tau
    buffers as tau_reduce writes them (per atom nlmsa^2, (L, m) order)."""
    L = _lib.lib()
    _sigs(L)
    if not tau:
        return tau
    na = max(len(nlmsa), 1)
    s = _lib_stream(stream)
    with Ordered(s):
        _lib.check(L.fdmgpu_tau_symmetrize(len(nlmsa), (C.c_int * na)(*[int(k) for k in nlmsa]),
                                           int(bool(ylm_comp)), len(tau), _dptrs([as_real(t) for t in tau]),
                                           _V(s.cuda_stream)))
    return tau
