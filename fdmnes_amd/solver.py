"""Host-side mirror of the reference's solver interface.

``mat_solve`` has the meaning of the reference's ``mat_solve``
(mat_solve_gaussian.f90:5-372) at the level of an assembled system: the
profile bounds ``lb1/lb2`` (1-based), the rows in compressed form (what the
Fortran shim gathers from ``calcMatRow``), and the right-hand sides, which
are replaced by the solution.  A vanishing pivot raises ``PivotError`` with
the reference's row number instead of ``STOP``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PivotError, check, ptr  # noqa: F401


def _prep(lb1, lb2, row_ptr, col, vals, rhs, cal_comp):
    lb1 = np.ascontiguousarray(lb1, np.int32)
    lb2 = np.ascontiguousarray(lb2, np.int32)
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    n, nl = rhs.shape
    if lb1.shape != (n,) or lb2.shape != (n,) or row_ptr.shape != (n + 1,):
        raise ValueError("inconsistent system sizes")
    are = np.ascontiguousarray(np.real(vals), np.float64)
    aim = np.ascontiguousarray(np.imag(vals), np.float64) if cal_comp else None
    smr = np.ascontiguousarray(np.real(rhs), np.float64).copy()
    smi = np.ascontiguousarray(np.imag(rhs), np.float64).copy() if cal_comp else None
    return n, nl, lb1, lb2, row_ptr, col, are, aim, smr, smi


def mat_solve(lb1, lb2, row_ptr, col, vals, rhs, cal_comp: bool):
    """FP64 drop-in solve.  rhs: (nligne, nlmso), complex when cal_comp.

    Returns (x, t_fill, t_tria)."""
    n, nl, lb1, lb2, row_ptr, col, are, aim, smr, smi = _prep(lb1, lb2, row_ptr, col, vals, rhs, cal_comp)
    L = _lib.lib()
    fr, tf, tt = C.c_int(0), C.c_double(0), C.c_double(0)
    rc = L.fdmgpu_mat_solve(n, nl, int(bool(cal_comp)), lb1, lb2, row_ptr, col, are, ptr(aim), smr, ptr(smi),
                            C.byref(fr), C.byref(tf), C.byref(tt))
    check(rc, fr.value)
    x = smr + 1j * smi if cal_comp else smr
    return x, tf.value, tt.value


def mat_solve_mixed(lb1, lb2, row_ptr, col, vals, rhs, cal_comp: bool, max_sweeps: int = 5, tol: float = 1e-12):
    """FP32 factors + FP64 refinement.  Returns (x, sweeps, final_rel);
    sweeps < 0 means the FP64 fallback was taken."""
    n, nl, lb1, lb2, row_ptr, col, are, aim, smr, smi = _prep(lb1, lb2, row_ptr, col, vals, rhs, cal_comp)
    L = _lib.lib()
    fr, sw, rel = C.c_int(0), C.c_int(0), C.c_double(0)
    rc = L.fdmgpu_mat_solve_mixed(n, nl, int(bool(cal_comp)), lb1, lb2, row_ptr, col, are, ptr(aim), smr, ptr(smi),
                                  int(max_sweeps), float(tol), C.byref(sw), C.byref(rel), C.byref(fr))
    check(rc, fr.value)
    x = smr + 1j * smi if cal_comp else smr
    return x, sw.value, rel.value


def finalize():
    """Releases the drop-in's kept device solvers (fdmgpu_finalize); the next
    call sets the library up again."""
    check(_lib.lib().fdmgpu_finalize())
