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


def _sigs(L):
    if getattr(L, "_tau_sigs", False):
        return
    P = C.POINTER
    L.fdmgpu_tau_reduce.argtypes = [P(TauDesc), C.c_int, P(_V), P(_V), P(_V), C.c_int, _V, _V]
    L._tau_sigs = True


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
    global _stream
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
    if stream is not None:
        s = library_stream(stream)
    else:
        if _stream is None:  # the module's own stream for calls that name none
            _stream = library_stream(None)
        s = _stream
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
