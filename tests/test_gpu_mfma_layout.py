"""In-bounds check of the 16x16x4 MFMA builtins' lane layout, which the tile
kernels assume: lane l holds A[l%16][l/16] and B[l/16][l%16]; output v of
lane l is D[l/16 + 4v][l%16] for FP64 and D[4*(l/16) + v][l%16] for FP32."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("prec", [0, 1])
def test_mfma_output_layout(prec):
    import torch  # noqa: F401  (one HIP runtime for the process)
    lib = ctypes.CDLL(os.path.join(ROOT, "tests", "libfdmgpu_probe.so"))
    out = np.zeros(256)
    assert lib.fdmgpu_probe_mfma(prec, out.ctypes.data_as(ctypes.c_void_p)) == 0
    A = np.array([[i + 16 * k + 1 for k in range(4)] for i in range(16)], float)
    B = np.array([[(k + 1) * 1000.0 + j for j in range(16)] for k in range(4)], float)
    D = A @ B
    got = out.reshape(64, 4)
    found = {}
    for l in range(64):
        for v in range(4):
            hits = np.argwhere(np.isclose(D, got[l, v], rtol=1e-6, atol=0))
            found[(l, v)] = [tuple(h) for h in hits]
    print({k: found[k] for k in [(0, 0), (0, 1), (1, 0), (16, 0), (17, 3)]}, flush=True)
    for l in range(64):
        for v in range(4):
            row = l // 16 + 4 * v if prec == 0 else 4 * (l // 16) + v
            assert (row, l % 16) in found[(l, v)], (l, v, found[(l, v)])
