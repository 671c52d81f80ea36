"""The FP64 drop-in on small random profile matrices against numpy: n = 64
exercises the diagonal-tile UL and the substitutions only, larger n the
TRSM and trailing-update kernels too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _banded(n, bl, bu, seed, cplx=True):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n), complex if cplx else float)
    for i in range(n):
        lo, hi = max(0, i - bl), min(n, i + bu + 1)
        A[i, lo:hi] = rng.standard_normal(hi - lo) + (1j * rng.standard_normal(hi - lo) if cplx else 0)
        A[i, i] += 4 * (bl + bu + 1)
    lb1 = np.maximum(np.arange(n) - bl, 0) + 1
    lb2 = np.minimum(np.arange(n) + bu, n - 1) + 1
    rp = [0]
    col, val = [], []
    for i in range(n):
        c = np.arange(lb1[i] - 1, lb2[i])
        col += list(c + 1)
        val += list(A[i, c])
        rp.append(len(col))
    return A, lb1.astype(np.int32), lb2.astype(np.int32), np.array(rp, np.int64), np.array(col, np.int32), np.array(val)


@pytest.mark.parametrize("n,bl,bu,cplx", [(64, 63, 63, True), (64, 10, 20, False), (130, 40, 70, True),
                                          (300, 90, 150, True), (300, 90, 150, False), (700, 200, 130, True)])
def test_dense_banded(n, bl, bu, cplx):
    from fdmnes_amd import solver
    A, lb1, lb2, rp, col, val = _banded(n, bl, bu, n + bl, cplx)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((n, 3)) + (1j * rng.standard_normal((n, 3)) if cplx else 0)
    print(f"case n={n} bl={bl} bu={bu} cplx={cplx}", flush=True)
    x, _, _ = solver.mat_solve(lb1, lb2, rp, col, val, b, cplx)
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-12
