"""Loader for libfdmgpu.so, the HIP library.

The product path has no CPU fallback: if the library is missing or no HIP
device can be initialised, every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdmgpu.so")
HOST_LIB_PATH = os.path.join(HERE, "libfdmgpu_host.so")

_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
vp = C.c_void_p


class FdmGpuError(RuntimeError):
    pass


class PivotError(FdmGpuError):
    """Vanishing pivot: the reference's write_error + STOP
    (mat_solve_gaussian.f90:266-273, :289-295)."""

    def __init__(self, row: int):
        super().__init__(f"Division par zero dans mat: ii = {row}")
        self.row = row


# (name, restype, argtypes) of every exported symbol, mirroring include/fdmgpu.h
SIGNATURES = {
    "fdmgpu_init": (C.c_int, [C.c_int]),
    "fdmgpu_finalize": (C.c_int, []),
    "fdmgpu_last_error": (C.c_char_p, []),
    "fdmgpu_solver_name": (None, [C.c_char_p]),
    "fdmgpu_mat_solve": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip, lp, ip, dp, vp, dp, vp,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fdmgpu_mat_solve_mixed": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip, lp, ip, dp, vp, dp, vp, C.c_int,
                                         C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int)]),
}


def load(path: str = LIB_PATH):
    """Loads the library without touching the GPU (symbol check only).

    torch ships its own HIP runtime with the same soname as the system one;
    it is imported first so that one runtime serves both (two runtimes in a
    process cannot share the device)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(path):
            raise FdmGpuError(f"{path} is missing: build it with `python -m fdmnes_amd.build`")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_inited = False


def lib():
    """Loads the library and initialises device 0."""
    global _inited
    L = load()
    if not _inited:
        check(L.fdmgpu_init(int(os.environ.get("FDMGPU_DEVICE", "0"))))
        _inited = True
    return L


_host = None


def load_host():
    """The host-only library (synthetic cluster builder); no HIP inside."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise FdmGpuError(f"{HOST_LIB_PATH} is missing: build it with `python -m fdmnes_amd.build`")
        _host = C.CDLL(HOST_LIB_PATH)
    return _host


def check(rc: int, fail_row: int = 0):
    if rc == 0:
        return
    if rc == 1:
        raise PivotError(fail_row)
    msg = load().fdmgpu_last_error()
    raise FdmGpuError(f"fdmgpu error {rc}: {msg.decode() if msg else ''}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
