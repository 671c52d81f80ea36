"""The oracle pinned by mathematics: brute force, closed forms, special
cases and the pivot thresholds — each chosen so a dropped term, a wrong sign
or index, or a transposed operand fails."""
import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O

BOHR = 0.52917721067


def _csr_from_dense(A):
    n = A.shape[0]
    rp = [0]
    col, val = [], []
    lb1 = np.zeros(n, np.int32)
    lb2 = np.zeros(n, np.int32)
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        nz = np.union1d(nz, [i])
        col += list(nz + 1)
        val += list(A[i, nz])
        rp.append(len(col))
        lb1[i] = nz.min() + 1
        lb2[i] = nz.max() + 1
    for i in range(n - 2, -1, -1):  # newind's monotone sweep of lb1
        lb1[i] = min(lb1[i], lb1[i + 1])
    return lb1, lb2, np.array(rp, np.int64), np.array(col, np.int32), np.array(val)


@pytest.mark.parametrize("E_comp,cal,basereel", [(True, True, False), (False, False, True), (False, True, False)])
def test_cluster_solve_vs_dense_bruteforce(E_comp, cal, basereel):
    cl = O.Cluster(1.5 / BOHR, 0.25 / BOHR, 3)
    V = np.random.default_rng(3).uniform(-0.4, 0.4, cl.npoint) - 0.6
    E, Ei = 0.9, (0.05 if E_comp else 0.0)
    A, b = cl.assemble_dense(V, E, Ei, E, E_comp, cal, basereel)
    st, x, *_ = cl.solve(V, E, Ei, E, E_comp, cal, basereel)
    assert st == 0
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-11
    # the residual itself is small in the operator's own scale
    assert np.abs(A @ x - b).max() / (np.abs(A).max() * np.abs(x).max()) < 1e-13


def test_dirichlet_cube_closed_form():
    """4th-order Laplacian + constant V - E on an m^3 cube with zero boundary:
    the 3-D sine transform diagonalises it; eigenvalues
    c - sum_axes [(4/3)(2 - 2cos t) - (1/12)(2 - 2cos 2t)]/h^2 with the
    second-neighbour couplings across the boundary truncated to zero, which
    breaks exact diagonality — so the check uses the 2nd-order (3-point)
    stencil the reference also supports (iord = 2), whose sine eigenpairs are
    exact."""
    m, h = 7, 0.5
    n = m ** 3
    shift = 0.3 - 0.1j
    idx = np.arange(n).reshape(m, m, m)
    A = np.zeros((n, n), complex)
    for (i, j, k), p in np.ndenumerate(idx):
        A[p, p] = 6 / h ** 2 + shift
        for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
            q = (i + d[0], j + d[1], k + d[2])
            if all(0 <= c < m for c in q):
                A[p, idx[q]] = -1 / h ** 2
    # radial-like ordering is irrelevant for the closed form; use the natural order
    t = np.pi * np.arange(1, m + 1) / (m + 1)
    S = np.sqrt(2 / (m + 1)) * np.sin(np.outer(np.arange(1, m + 1), np.arange(1, m + 1)) * np.pi / (m + 1))
    lam1 = (2 - 2 * np.cos(t)) / h ** 2
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :] + shift
    rng = np.random.default_rng(0)
    b = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    B = b.reshape(m, m, m, 3)
    Bh = np.einsum("ai,bj,ck,ijkr->abcr", S, S, S, B)
    X = np.einsum("ai,bj,ck,ijkr->abcr", S, S, S, Bh / lam[..., None]).reshape(n, 3)
    lb1, lb2, rp, col, val = _csr_from_dense(A)
    st, x, _ = O.mat_solve_csr(lb1, lb2, rp, col, val, b, True)
    assert st == 0
    assert np.abs(x - X).max() / np.abs(X).max() < 1e-12


def test_pentadiagonal_thomas_and_structure():
    """1-D 4th-order operator (pentadiagonal): solved against a dense solve and
    its transpose — a transposed operand in the elimination fails this since
    the matrix is made non-symmetric."""
    n = 60
    rng = np.random.default_rng(1)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 7.5 + rng.uniform(-0.2, 0.2)
        for d, c in [(1, -4 / 3), (2, 1 / 12)]:
            if i + d < n:
                A[i, i + d] = c * 1.3
            if i - d >= 0:
                A[i, i - d] = c * 0.7
    b = rng.standard_normal((n, 2))
    lb1, lb2, rp, col, val = _csr_from_dense(A)
    st, x, _ = O.mat_solve_csr(lb1, lb2, rp, col, val, b, False)
    assert st == 0
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-13
    assert np.abs(x - np.linalg.solve(A.T, b)).max() > 1e-3


def test_diagonal_and_single_row():
    A = np.diag([2.0, -4.0, 0.5])
    lb1, lb2, rp, col, val = _csr_from_dense(A)
    st, x, _ = O.mat_solve_csr(lb1, lb2, rp, col, val, np.ones((3, 1)), False)
    assert st == 0 and np.allclose(x[:, 0], [0.5, -0.25, 2.0], rtol=0, atol=0)
    A1 = np.array([[3.0 - 4.0j]])
    lb1, lb2, rp, col, val = _csr_from_dense(A1)
    st, x, _ = O.mat_solve_csr(lb1, lb2, rp, col, val, np.array([[1.0 + 0j]]), True)
    assert st == 0 and abs(x[0, 0] - 1 / (3 - 4j)) < 1e-16


@pytest.mark.parametrize("cal", [True, False])
def test_pivot_thresholds(cal):
    """The reference stops at |a|^2 < 1e-20 (complex) and |a| < 1e-10 (real);
    the elimination is bottom-up, so the reported row is the last failing one."""
    for d, fails in [(0.9e-10, True), (1.1e-10, False)]:
        A = np.diag([1.0, d, 2.0, 1.0]).astype(complex if cal else float)
        A[1, 2] = 0.5  # upper coupling eliminated before the pivot of row 2
        lb1, lb2, rp, col, val = _csr_from_dense(A)
        st, x, fr = O.mat_solve_csr(lb1, lb2, rp, col, val, np.ones((4, 1), complex if cal else float), cal)
        assert (st == 1) == fails
        if fails:
            assert fr == 2


def test_bessel_closed_forms_and_scipy():
    for z in [0.37, 1.2, 3.3, 8.9]:
        b, n = O.cbessneur(1.0, z, 6)
        assert abs(b[0] - np.sin(z) / z) < 1e-15 and abs(n[0] + np.cos(z) / z) < 1e-15
        assert abs(b[1] - (np.sin(z) / z ** 2 - np.cos(z) / z)) < 1e-15
        for L in range(7):
            assert abs(n[L] - sp.spherical_yn(L, z)) <= 1e-11 * abs(sp.spherical_yn(L, z))
            if z > L:  # upward recurrence of j_l is accurate above the turning point
                assert abs(b[L] - sp.spherical_jn(L, z)) <= 1e-10 * max(abs(sp.spherical_jn(L, z)), 1e-3)
    zc = 2.5 + 0.3j
    f = 0.7 + 0.1j
    b, n = O.cbessneu(f, zc, 3)
    assert abs(b[0] - f * np.sin(zc) / zc) < 1e-15 and abs(n[0] + f * np.cos(zc) / zc) < 1e-15
    for L in range(4):
        assert abs(b[L] - f * sp.spherical_jn(L, zc)) < 1e-12
        assert abs(n[L] - f * sp.spherical_yn(L, zc)) < 1e-12


def test_real_ylm_against_scipy_and_orthonormality():
    rng = np.random.default_rng(2)
    v = rng.standard_normal((400, 3))
    lmax = 4
    Y = np.array([O.ylm_real(lmax, p, float(np.linalg.norm(p))) for p in v])
    r = np.linalg.norm(v, axis=1)
    theta = np.arccos(v[:, 2] / r)
    phi = np.arctan2(v[:, 1], v[:, 0])
    for L in range(lmax + 1):
        for m in range(-L, L + 1):
            if hasattr(sp, "sph_harm_y"):
                yc = sp.sph_harm_y(L, abs(m), theta, phi)
            else:
                yc = sp.sph_harm(abs(m), L, phi, theta)
            if m == 0:
                ref = yc.real
            elif m > 0:
                ref = np.sqrt(2) * (-1) ** m * yc.real
            else:
                ref = np.sqrt(2) * (-1) ** m * yc.imag
            assert np.abs(Y[:, L * L + L + m] - ref).max() < 1e-13, (L, m)
    # orthonormality on a fine quadrature of the sphere
    th = np.linspace(0, np.pi, 201)
    ph = np.linspace(0, 2 * np.pi, 401)[:-1]
    TT, PP = np.meshgrid(th, ph, indexing="ij")
    pts = np.stack([np.sin(TT) * np.cos(PP), np.sin(TT) * np.sin(PP), np.cos(TT)], -1).reshape(-1, 3)
    Yg = np.array([O.ylm_real(2, p, 1.0) for p in pts])
    w = (np.sin(TT) * (th[1] - th[0]) * (ph[1] - ph[0])).reshape(-1)
    G = (Yg * w[:, None]).T @ Yg
    assert np.abs(G - np.eye(9)).max() < 2e-3


def test_profile_invariants():
    cl = O.Cluster(2.0 / BOHR, 0.25 / BOHR, 3)
    lb1, lb2, newinv, newp = cl.profile()
    ii = np.arange(1, cl.nligne + 1)
    assert np.all(np.diff(lb1) >= 0)
    assert np.all(lb1 <= ii) and np.all(lb2 >= ii) and np.all(lb2 <= cl.nligne)
    assert np.all(newinv[: cl.npoint] == ii[: cl.npoint]) and np.all(newinv[cl.npoint:] == 0)
    # boundary rows couple to the outer unknowns: lb2 = nligne exactly for them
    isrt, _ = cl.boundary
    assert np.all(lb2[cl.npoint:] == cl.nligne)
    # structural pattern lies inside the profile (fill closure relies on it)
    V = np.zeros(cl.npoint)
    rp, col, vals, _ = cl.assemble_csr(V, 0.5, 0.02, 0.5, True, True, False)
    rows = np.repeat(np.arange(cl.nligne), np.diff(rp))
    assert np.all(col >= lb1[rows]) and np.all(col <= lb2[rows])


def test_grid_order_is_radial_with_reference_ties():
    cl = O.Cluster(1.5 / BOHR, 0.25 / BOHR, 2)
    r = cl.xyz[:, 3]
    assert np.all(np.diff(r) > -1e-4)
    # the sphere holds exactly the points with |index| <= R/h (+ epspos)
    ind = cl.indice
    d = np.sqrt((ind ** 2).sum(1))
    assert np.all(d[: cl.npoint] * (0.25 / BOHR) <= 1.5 / BOHR + 1e-4)
