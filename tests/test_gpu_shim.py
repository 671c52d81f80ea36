"""The Fortran drop-in (fdmnes_amd/fortran/mat_solve_hip.f90) behind the
reference's own 93-argument mat_solve surface: the reference's calcMatRow
assembles the rows, the shim solves on the GPU, and the result is compared
with the reference's Gaussian mat_solve on the same inputs."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "fdmnes_amd", "libfdmnes_shim.so")
BOHR = 0.52917721067


@pytest.mark.parametrize("E_comp,cal,basereel", [(True, True, False), (False, False, True), (False, True, False)])
def test_shim_matches_reference_mat_solve(E_comp, cal, basereel):
    from oracle import oracle as O
    from oracle import ref as R
    if not (R.available() and os.path.exists(SHIM)):
        pytest.skip("needs oracle/_ref and the shim (built where the reference sources are)")
    import torch  # noqa: F401  (one HIP runtime for the process)
    from fdmnes_amd import _lib
    _lib.lib()
    R.lib()
    shim = ctypes.CDLL(SHIM, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    cl = O.Cluster(1.5 / BOHR, 0.25 / BOHR, 3)
    V = np.random.default_rng(7).uniform(-0.5, 0.5, cl.npoint) - 0.9
    E, Ei = 1.0, (0.04 if E_comp else 0.0)
    bes, neu = R.phiso(cl, E_comp, E, Ei, E)
    print(f"case n={cl.nligne} E_comp={E_comp} cal={cal} basereel={basereel}", flush=True)
    x_ref = R.mat_solve(cl, V, E, Ei, bes, neu, E_comp, cal, basereel)
    x_gpu = R.mat_solve(cl, V, E, Ei, bes, neu, E_comp, cal, basereel, fn=shim.mat_solve_)
    rel = np.abs(x_gpu - x_ref).max() / np.abs(x_ref).max()
    assert rel <= 1e-10, rel
    name = ctypes.create_string_buffer(5)
    a, b = ctypes.c_int(0), ctypes.c_int(0)
    shim.getsolverparams_(ctypes.byref(a), ctypes.byref(b), name, ctypes.c_long(5))
    assert name.raw == b"HIP  " and a.value == 1
