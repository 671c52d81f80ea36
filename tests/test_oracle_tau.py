"""The oracle's post-solve reduction (invcomp and the contraction to the atom
amplitude products, mat.f90:297-405 and :1927-1987) pinned by the compiled
reference, by a physics identity between the real and the complex outer-sphere
bases, and by structural invariants."""
import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067
SIZES = (1, 2, 3, 9, 16, 25, 49, 64, 81)


def _rand(rng, n):
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    if n > 1:
        A[0, 0] = 0  # the first step needs a row interchange
    return A


@pytest.mark.ref
def test_invcomp_matches_reference_bit_for_bit():
    from oracle import ref as R
    if not R.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(0)
    for n in SIZES:
        A = _rand(rng, n)
        inv, info = O.invcomp(A)
        assert info == 0
        assert np.array_equal(inv, R.invcomp(A)), n


def test_invcomp_is_the_inverse():
    rng = np.random.default_rng(1)
    for n in SIZES:
        A = _rand(rng, n)
        inv, info = O.invcomp(A)
        assert info == 0
        assert np.abs(inv @ A - np.eye(n)).max() <= 1e-12
        assert np.abs(inv - np.linalg.inv(A)).max() <= 1e-13 * np.abs(inv).max()


def test_invcomp_reports_the_first_zero_pivot():
    A = np.eye(4, dtype=complex)
    A[:, 2] = 0  # column 3 vanishes below the pivot search: ZGETF2 INFO = 3
    _, info = O.invcomp(A)
    assert info == 3


def _solutions(cl, V, E, Ec):
    """Regular solutions of one real energy in the real (Neumann) and the
    complex (Bessel + i Neumann) outer-sphere bases."""
    st_r, xr, *_ = cl.solve(V, E, 0.0, Ec, False, False, True)
    st_c, xc, *_ = cl.solve(V, E, 0.0, Ec, False, True, False)
    assert st_r == 0 and st_c == 0
    return xr, xc


@pytest.mark.parametrize("fdm_comp", [False, True])
def test_real_basis_renormalised_equals_complex_basis(fdm_comp):
    """With outgoing functions h = j + i n the complex-basis solution of RHS
    L is the real-basis combination sum_M x_M C_ML, C = (I + i X)^-1, X the
    real solutions' outer-sphere amplitudes, on every grid row: exactly the
    renormalisation mat.f90:312-343 applies before the contraction.  So tau of
    the real basis (basereel) must equal tau of the complex basis without
    renormalisation; a transposed norm, a dropped identity or a wrong sign of
    i breaks it."""
    cl = O.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    rng = np.random.default_rng(2)
    V = rng.uniform(-1.5, 0.0, cl.npoint)
    xr, xc = _solutions(cl, V, 0.3, 0.6)
    atoms = [(0, 9), (37, 4), (cl.nligneso - 5, 5)]
    tr, norm, info = O.tau(xr, cl.nligneso, True, fdm_comp, atoms, want_norm=True)
    tc, _, info_c = O.tau(xc, cl.nligneso, False, fdm_comp, atoms)
    assert info == 0 and info_c == 0
    for a, b in zip(tr, tc):
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()
    # the identity separates the orientations: the amplitudes norm^-1 x
    # (norm_t contracted over its first index) match, those of the
    # transposed normalisation do not
    Xm = xr[cl.nligneso:, :]  # Xm[L', M]: outer amplitude L' of real solution M
    wrong = np.linalg.inv(np.eye(cl.nlmso) + 1j * Xm) @ xr[0:9].T
    good = norm @ xr[0:9].T
    assert np.abs(good.T - xc[0:9]).max() <= 1e-11 * np.abs(xc).max()
    assert np.abs(wrong.T - xc[0:9]).max() > 1e-6 * np.abs(xc).max()


def test_tau_without_outer_amplitudes_is_the_plain_contraction():
    """X = 0 makes norm the identity: the inverse is exactly I and the
    renormalised tau equals the unrenormalised one bit for bit."""
    rng = np.random.default_rng(3)
    n, nl = 40, 9
    x = rng.normal(size=(n, nl))
    x[n - nl:, :] = 0
    atoms = [(0, 6), (10, 9)]
    t1, norm, info = O.tau(x, n - nl, True, False, atoms, want_norm=True)
    t0, _, _ = O.tau(x, n - nl, False, False, atoms)
    assert info == 0 and np.array_equal(norm, np.eye(nl, dtype=complex))
    for a, b in zip(t1, t0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("cal", [False, True])
def test_tau_structure(cal):
    """-i a^T conj(a): i tau is Hermitian positive semi-definite (-Im of its
    diagonal is a density of states); -i a^T a (FDM_comp) is symmetric;
    a second call accumulates."""
    rng = np.random.default_rng(4)
    n, nl = 50, 16
    x = rng.normal(size=(n, nl)) + (1j * rng.normal(size=(n, nl)) if cal else 0)
    atoms = [(3, 16), (30, 7)]
    t, _, info = O.tau(x, n - nl, True, False, atoms)
    assert info == 0
    for a in t:
        h = 1j * a
        assert np.abs(h - h.conj().T).max() <= 1e-12 * np.abs(h).max()
        assert np.linalg.eigvalsh(0.5 * (h + h.conj().T)).min() >= -1e-10 * np.abs(h).max()
        assert (-a.diagonal().imag >= 0).all()
    ts, _, _ = O.tau(x, n - nl, True, True, atoms)
    for a in ts:
        assert np.abs(a - a.T).max() <= 1e-12 * np.abs(a).max()
    t2, _, _ = O.tau(x, n - nl, True, False, atoms, tau0=t)
    for a, b in zip(t2, t):
        assert np.abs(a - 2 * b).max() <= 1e-13 * np.abs(b).max()


def test_tau_reports_a_singular_normalisation():
    """norm = I + i X with X = diag(1, i, 1, 1): its second diagonal entry is
    1 + i*i = 0, and the reference stops at ZGETF2's INFO = 2."""
    n, nl = 20, 4
    x = np.zeros((n, nl), complex)
    x[n - nl:, :] = np.diag([1, 1j, 1, 1])
    _, _, info = O.tau(x, n - nl, True, False, [(0, 2)])
    assert info == 2
