"""Device-resident batched solver (``fdmgpu_solver_*`` in include/fdmgpu.h).

Device memory comes from torch tensors.  The library launches on a stream of
its own (graph capture needs one; torch's default stream is the legacy null
stream); every call is ordered after the work already queued on the caller's
current torch stream and before the work queued there afterwards, so torch
events on the caller's stream bracket the library's kernels."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

FP64, FP32 = 0, 1
_V = C.c_void_p


def _sigs(L):
    if getattr(L, "_solver_sigs", False):
        return
    P = C.POINTER
    L.fdmgpu_solver_create.argtypes = [C.c_int, C.c_int, C.c_int, _V, P(_V)]
    L.fdmgpu_solver_destroy.argtypes = [_V]
    L.fdmgpu_solver_destroy.restype = None
    L.fdmgpu_solver_add_system.argtypes = [_V, C.c_int, _lib.ip, _lib.ip, _lib.lp, _lib.ip, P(C.c_int)]
    L.fdmgpu_solver_solve.argtypes = [_V, C.c_int, P(C.c_int), P(_V), P(_V), P(_V), C.c_int, C.c_double,
                                      P(C.c_int), P(C.c_int), P(C.c_double)]
    L.fdmgpu_solver_adjoint.argtypes = [_V, C.c_int, P(C.c_int), P(_V), P(_V), P(_V)]
    L.fdmgpu_solver_sigma.argtypes = [_V, C.c_int, P(C.c_int), P(_V), P(_V), P(C.c_double)]
    L.fdmgpu_solver_sigma_noalloc.argtypes = [_V, C.c_int, P(C.c_int), P(_V), P(_V), _V]
    L.fdmgpu_solver_options.argtypes = [_V, C.c_int, C.c_int, C.c_int]
    L.fdmgpu_solver_stats.argtypes = [_V, P(C.c_double)]
    L.fdmgpu_solver_tune.argtypes = [_V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.fdmgpu_solver_overlap.argtypes = [_V, C.c_int]
    L.fdmgpu_solver_snapshots.argtypes = [_V, C.c_int]
    L.fdmgpu_solver_snapshot.argtypes = [_V, C.c_int, C.c_int, C.c_int, _V, C.c_int64, P(C.c_int64)]
    L._solver_sigs = True


def _ptrs(ts):
    arr = (_V * len(ts))()
    for k, t in enumerate(ts):
        if not t.is_contiguous():
            raise ValueError("device buffers must be contiguous")
        arr[k] = t.data_ptr()
    return arr


def library_stream(stream: torch.cuda.Stream | None) -> torch.cuda.Stream:
    """The stream a library object launches on: the given one unless it is
    the null stream, else a new one (default priority)."""
    if stream is not None and stream.cuda_stream != 0:
        return stream
    return torch.cuda.Stream()


class Ordered:
    """Context of one library call on stream s: s waits for the caller's
    current stream first, and the caller's stream waits for s afterwards."""

    def __init__(self, s: torch.cuda.Stream):
        self.s = s

    def __enter__(self):
        self.cur = torch.cuda.current_stream()
        if self.cur != self.s:
            self.s.wait_stream(self.cur)
        return self

    def __exit__(self, *exc):
        if self.cur != self.s:
            self.cur.wait_stream(self.s)
        return False


def as_real(t: torch.Tensor) -> torch.Tensor:
    """Interleaved float64 view of a complex128 tensor (float64 passes through)."""
    return torch.view_as_real(t) if t.is_complex() else t


class DeviceSolver:
    def __init__(self, precision: int, cal_comp: bool, nrhs: int, stream: torch.cuda.Stream | None = None):
        L = _lib.lib()
        _sigs(L)
        self.L = L
        self.cal = bool(cal_comp)
        self.nrhs = nrhs
        self.precision = precision
        self.stream = library_stream(stream)
        h = _V()
        _lib.check(L.fdmgpu_solver_create(precision, int(self.cal), nrhs, _V(self.stream.cuda_stream), C.byref(h)))
        self._h = h
        self.sizes = []

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.fdmgpu_solver_destroy(self._h)
            self._h = None

    def add_system(self, lb1, lb2, row_ptr, col) -> int:
        s = C.c_int(-1)
        _lib.check(self.L.fdmgpu_solver_add_system(self._h, len(lb1), np.ascontiguousarray(lb1, np.int32),
                                                   np.ascontiguousarray(lb2, np.int32),
                                                   np.ascontiguousarray(row_ptr, np.int64),
                                                   np.ascontiguousarray(col, np.int32), C.byref(s)))
        self.sizes.append(len(lb1))
        return s.value

    def options(self, use_graphs=True, time_updates=False, persist_wg=0):
        _lib.check(self.L.fdmgpu_solver_options(self._h, int(use_graphs), int(time_updates), int(persist_wg)))

    def overlap(self, on: bool):
        """The left-looking update of the next block row and column on a
        second stream beside the current step's panel (fdmgpu_solver_overlap)."""
        _lib.check(self.L.fdmgpu_solver_overlap(self._h, int(bool(on))))

    def tune(self, gk=-1, trail_wg=0, persist_solve=-1, solve_wg=0, trail_tile=0, lookahead=-1):
        _lib.check(self.L.fdmgpu_solver_tune(self._h, int(gk), int(trail_wg), int(persist_solve), int(solve_wg),
                                             int(trail_tile), int(lookahead)))

    def solve(self, slots, vals, b, x, max_sweeps=5, tol=1e-12, raise_on_pivot=True):
        """vals/b/x: lists of device tensors (complex128 or float64)."""
        n = len(slots)
        ids = (C.c_int * n)(*slots)
        fr = (C.c_int * n)()
        sw = (C.c_int * n)()
        rel = (C.c_double * n)()
        with Ordered(self.stream):
            rc = self.L.fdmgpu_solver_solve(self._h, n, ids, _ptrs([as_real(v) for v in vals]),
                                            _ptrs([as_real(t) for t in b]), _ptrs([as_real(t) for t in x]),
                                            int(max_sweeps), float(tol), fr, sw, rel)
        if rc == 1 and not raise_on_pivot:
            return list(fr), list(sw), list(rel)
        _lib.check(rc, max(fr))
        return list(fr), list(sw), list(rel)

    def adjoint(self, slots, src, x, grad):
        n = len(slots)
        ids = (C.c_int * n)(*slots)
        with Ordered(self.stream):
            rc = self.L.fdmgpu_solver_adjoint(self._h, n, ids, _ptrs([as_real(t) for t in src]),
                                              _ptrs([as_real(t) for t in x]), _ptrs(grad))
        _lib.check(rc)

    def sigma(self, slots, src, x) -> np.ndarray:
        n = len(slots)
        ids = (C.c_int * n)(*slots)
        out = (C.c_double * n)()
        with Ordered(self.stream):
            rc = self.L.fdmgpu_solver_sigma(self._h, n, ids, _ptrs([as_real(t) for t in src]),
                                            _ptrs([as_real(t) for t in x]), out)
        _lib.check(rc)
        return np.array(out[:])

    def sigma_noalloc(self, slots, src, x, out: torch.Tensor):
        """sigma into the float64 device tensor out (len(slots) entries),
        allocating nothing (fdmgpu_solver_sigma_noalloc)."""
        if out.dtype != torch.float64 or not out.is_cuda or out.numel() < len(slots) or not out.is_contiguous():
            raise ValueError("out: contiguous float64 device tensor with one entry per point")
        n = len(slots)
        ids = (C.c_int * n)(*slots)
        with Ordered(self.stream):
            rc = self.L.fdmgpu_solver_sigma_noalloc(self._h, n, ids, _ptrs([as_real(t) for t in src]),
                                                    _ptrs([as_real(t) for t in x]), _V(out.data_ptr()))
        _lib.check(rc)

    def snapshots(self, on: bool):
        """Keep host copies of each stage of the next solves (test hook,
        fdmgpu_solver_snapshots)."""
        _lib.check(self.L.fdmgpu_solver_snapshots(self._h, int(bool(on))))

    def snapshot(self, slot: int) -> dict:
        """The last solve's stages of one slot: raw bytes of the factor and
        inverse tiles, and per refinement sweep of the FP32 correction and
        the FP64 x after it."""
        def get(what, sweep=0):
            n = C.c_int64(0)
            _lib.check(self.L.fdmgpu_solver_snapshot(self._h, slot, what, sweep, None, 0, C.byref(n)))
            if what == 3:
                return n.value
            buf = (C.c_char * n.value)()
            _lib.check(self.L.fdmgpu_solver_snapshot(self._h, slot, what, sweep, buf, n.value, C.byref(n)))
            return bytes(buf)
        nsw = get(3)
        return {"factor": get(0), "d": [get(1, k) for k in range(nsw)], "x": [get(2, k) for k in range(nsw)]}

    def stats(self) -> dict:
        out = (C.c_double * 8)()
        _lib.check(self.L.fdmgpu_solver_stats(self._h, out))
        return {"update_ms": out[0], "update_launches": int(out[1]), "tile_macs": out[2],
                "profile_macs": out[3], "factor_bytes": out[4], "profile_entries": out[5],
                "tile_other_macs": out[6], "update_bytes": out[7]}
