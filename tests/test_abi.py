"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fdmnes_amd", "libfdmgpu.so")
HOST_LIB = os.path.join(ROOT, "fdmnes_amd", "libfdmgpu_host.so")
# header -> the library that exports it
LIBS = {"fdmgpu.h": LIB, "fdmgpu_cluster.h": HOST_LIB}


def _declared(header=None):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if header and os.path.basename(h) != header:
            continue
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fdmgpu_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "fdmgpu_mat_solve" in names and "fdmgpu_solver_adjoint" in names, names


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_symbol(header):
    path = LIBS[header]
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    lib = ctypes.CDLL(path)
    names = _declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_header_has_a_library():
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        assert os.path.basename(h) in LIBS


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfdmgpu.so not built")
def test_solver_name_without_gpu():
    lib = ctypes.CDLL(LIB)
    buf = ctypes.create_string_buffer(6)
    lib.fdmgpu_solver_name(buf)
    assert buf.raw[:5] == b"HIP  "


def test_python_signatures_cover_header():
    from fdmnes_amd import _lib
    for n in ["fdmgpu_init", "fdmgpu_mat_solve", "fdmgpu_mat_solve_mixed", "fdmgpu_last_error"]:
        assert n in _lib.SIGNATURES


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfdmgpu.so not built")
def test_entry_points_fail_loudly_without_gpu():
    """Without a device the library reports an error status with a message
    (no crash), and the Python drop-in raises instead of computing anything
    on the host."""
    import numpy as np
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the device path is tested by the GPU tests")
    lib = ctypes.CDLL(LIB)
    lib.fdmgpu_last_error.restype = ctypes.c_char_p
    # argument checks come before any device call
    rc = lib.fdmgpu_mat_solve(0, 1, 1, None, None, None, None, None, None, None, None, None, None, None)
    assert rc == -1  # FDMGPU_EINVAL
    assert b"invalid argument" in lib.fdmgpu_last_error()
    rc = lib.fdmgpu_init(0)
    assert rc < 0 and lib.fdmgpu_last_error()
    from fdmnes_amd import _lib, solver
    n, nrhs = 3, 1
    lb1 = np.array([1, 1, 2], np.int32)
    lb2 = np.array([2, 3, 3], np.int32)
    row_ptr = np.array([0, 2, 5, 7], np.int64)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    vals = np.array([4, 1, 1, 4, 1, 1, 4], np.complex128)
    rhs = np.ones((n, nrhs), np.complex128)
    with pytest.raises(_lib.FdmGpuError, match="device"):
        solver.mat_solve(lb1, lb2, row_ptr, col, vals, rhs, True)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfdmgpu.so not built")
def test_tau_reduce_argument_checks_without_gpu():
    """fdmgpu_tau_reduce rejects, before any device call, an nlmso outside
    1..256, atom rows outside the system and blocks whose normalisation,
    amplitudes and pivots would not fit one workgroup's 160 KiB of LDS."""
    from fdmnes_amd.tau import TauDesc
    lib = ctypes.CDLL(LIB)
    lib.fdmgpu_last_error.restype = ctypes.c_char_p
    dummy = (ctypes.c_void_p * 1)(1)  # never dereferenced: every case fails its checks first

    def call(nligne, nlmso, atoms, count=1):
        ianew = (ctypes.c_int * max(len(atoms), 1))(*[a for a, _ in atoms])
        nlmsa = (ctypes.c_int * max(len(atoms), 1))(*[k for _, k in atoms])
        d = TauDesc(nligne, nlmso, nligne - nlmso, 1, 1, 0, len(atoms), ianew, nlmsa)
        rc = lib.fdmgpu_tau_reduce(ctypes.byref(d), count, dummy, dummy, None, 0, None, None)
        return rc, lib.fdmgpu_last_error().decode()

    rc, msg = call(1000, 300, [(0, 9)])
    assert rc == -1 and "nlmso <= 256" in msg
    rc, msg = call(1000, 16, [(995, 9)])
    assert rc == -1 and "rows of atom 0" in msg
    # 128^2 + 128 * (128 + 1) complex entries: 526 KiB of LDS
    rc, msg = call(4000, 128, [(0, 128)])
    assert rc == -1 and "160 KiB" in msg
    # a block that fits (nlmso 64, atoms of 64 rows: 129 KiB); no points, so no device call
    rc, msg = call(4000, 64, [(0, 64)], count=0)
    assert rc == 0
