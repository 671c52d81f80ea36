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
