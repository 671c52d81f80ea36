"""A tile-level model of the GPU factorisation's launch sequence (block
skyline of the solver's plan, panels, L-row / U-column products, one-step
updates or left-looking updates plus one group rectangle per gk steps, and the
per-step overlap that moves part of each left-looking update to a second
stream), run in numpy on random skyline matrices with 16 x 16 tiles:

- the tiles end as U and L with A = U L (the reference's factorisation
  without pivoting, mat_solve_gaussian.f90:196-302) and every product the
  schedule forms lands in a stored tile;
- the overlap gives the same factors, and the tiles the second-stream launch
  reads and writes are disjoint from those the main stream writes and reads
  while it runs (it is joined before the next left-looking launch);
- so does the group-level lookahead, whose remainder of each group's
  rectangle runs on the second stream beside the next group's chain."""
import numpy as np
import pytest

TB = 16


def _plan(lb1, lb2):
    n = len(lb1)
    NB = (n + TB - 1) // TB
    J0, J1 = np.zeros(NB, int), np.zeros(NB, int)
    for I in range(NB):
        r0, r1 = I * TB, min(n, (I + 1) * TB)
        J0[I] = min(I, (lb1[r0] - 1) // TB)
        J1[I] = max(I, max((lb2[r] - 1) // TB for r in range(r0, r1)))
    J1 = np.maximum.accumulate(J1)
    I0 = np.zeros(NB, int)
    i = 0
    for K in range(NB):
        while J1[i] < K:
            i += 1
        I0[K] = i
    return NB, J0, J1, I0


def _matrix(n, seed, low_min=2, low_max=45, up_max=70, diag=30.0):
    rng = np.random.default_rng(seed)
    low = rng.integers(low_min, low_max, n)
    lb1 = np.maximum.accumulate(np.maximum(1, np.arange(1, n + 1) - low))
    lb2 = np.minimum(n, np.arange(1, n + 1) + rng.integers(0, up_max, n))
    A = np.zeros((n, n), complex)
    for i in range(n):
        k = np.arange(lb1[i] - 1, lb2[i])
        A[i, k] = rng.normal(size=len(k)) + 1j * rng.normal(size=len(k))
        A[i, i] += diag  # no pivoting: keep every pivot away from zero
    return A, lb1, lb2


class Model:
    def __init__(self, A, lb1, lb2):
        n = len(lb1)
        self.NB, self.J0, self.J1, self.I0 = _plan(lb1, lb2)
        N = self.NB * TB
        Ap = np.eye(N, dtype=complex)
        Ap[:n, :n] = A
        self.A = Ap
        self.T = {}
        for I in range(self.NB):
            for J in range(self.J0[I], self.J1[I] + 1):
                self.T[I, J] = Ap[I * TB:(I + 1) * TB, J * TB:(J + 1) * TB].copy()
        # every entry of A lies in a stored tile
        for i in range(n):
            for j in np.nonzero(A[i])[0]:
                assert (i // TB, j // TB) in self.T

    # each launch returns (reads, writes) as sets of tiles
    def panel(self, K):
        D = self.T[K, K]
        for q in range(TB - 1, -1, -1):
            D[q, :q] /= D[q, q]
            D[:q, :q] -= np.outer(D[:q, q], D[q, :q])
        return {(K, K)}, {(K, K)}

    def products(self, K):
        D = self.T[K, K]
        U = np.triu(D)
        L = np.tril(D, -1) + np.eye(TB)
        r, w = {(K, K)}, set()
        for J in range(self.J0[K], K):
            self.T[K, J] = np.linalg.solve(U, self.T[K, J])
            r.add((K, J))
            w.add((K, J))
        for I in range(self.I0[K], K):
            self.T[I, K] = np.linalg.solve(L.T, self.T[I, K].T).T
            r.add((I, K))
            w.add((I, K))
        return r, w

    def update(self, Kb, nk, i0, ni, j0, nj):
        """k_trail item: tiles [i0, i0+ni) x [j0, j0+nj) accumulate the steps
        Kb..Kb+nk-1 that reach them (a prefix: I0 and J0 are non-decreasing)."""
        r, w = set(), set()
        for I in range(i0, i0 + ni):
            for J in range(j0, j0 + nj):
                assert (I, J) in self.T, "a product outside the stored tiles"
                for K2 in range(Kb, Kb + nk):
                    if not (self.I0[K2] <= I and self.J0[K2] <= J):
                        break
                    self.T[I, J] = self.T[I, J] - self.T[I, K2] @ self.T[K2, J]
                    r |= {(I, K2), (K2, J)}
                r.add((I, J))
                w.add((I, J))
        return r, w

    def factor(self, gk, overlap=False, side_first=False, lookahead=False):
        """factor()'s sequence; with the overlap, the second-stream launch of
        each step is checked against the main stream's launches it runs
        beside and applied before or after them (side_first).  With the
        lookahead a group's rectangle is split: the next group's block rows
        and columns on the main stream, the rest on the second stream from
        then until the main stream waits for it before the next group's
        rectangle (checked against every main-stream launch of that window,
        applied at its start or at its end).  Returns the number of tiles the
        remainders updated."""
        NB, J0, I0 = self.NB, self.J0, self.I0
        pending = None  # remainder on the second stream: [launch, reads, writes, result or None]
        rest_tiles = 0

        def join():
            nonlocal pending
            launch, R, W, res = pending
            r, w = res if res is not None else launch()
            assert not (w & (R | W)), "the remainder writes a tile the main stream uses"
            assert not (r & W), "the remainder reads a tile the main stream writes"
            pending = None

        for t in range(NB):
            K = NB - 1 - t
            nJ, nI = K - J0[K], K - I0[K]
            stream = []  # this step's main-stream launches, in order
            side = None
            new_rest = None
            if gk <= 0:
                stream += [lambda K=K: self.panel(K), lambda K=K: self.products(K)]
                if nI > 0 and nJ > 0:
                    stream.append(lambda K=K: self.update(K, 1, I0[K], nI, J0[K], nJ))
            else:
                Khi = NB - 1 - gk * (t // gk)
                Klo = max(Khi - gk + 1, 0)
                if K < Khi:
                    nk = 1 if overlap else Khi - K
                    stream.append(lambda K=K, nk=nk: self.update(K + 1, nk, K, 1, J0[K], nJ + 1))
                    stream.append(lambda K=K, nk=nk: self.update(K + 1, nk, I0[K], nI, K, 1))
                if overlap and Klo < K < Khi:
                    K2 = K - 1
                    nJ2, nI2 = K2 - J0[K2], K2 - I0[K2]

                    def side(K=K, K2=K2, nJ2=nJ2, nI2=nI2, Khi=Khi):
                        r1, w1 = self.update(K + 1, Khi - K, K2, 1, J0[K2], nJ2 + 1)
                        r2, w2 = self.update(K + 1, Khi - K, I0[K2], nI2, K2, 1)
                        return r1 | r2, w1 | w2
                stream += [lambda K=K: self.panel(K), lambda K=K: self.products(K)]
                if K == Klo and Klo > 0:
                    nk, i0, j0 = Khi - Klo + 1, I0[Klo], J0[Klo]
                    stream.append("join")  # the main stream waits for the earlier remainder
                    if lookahead:
                        Klo2 = max(Klo - gk, 0)  # the next group's steps are Klo2..Klo-1
                        a, b = max(i0, Klo2), max(j0, Klo2)
                        stream.append(lambda: self.update(Klo, nk, a, Klo - a, j0, Klo - j0))
                        stream.append(lambda: self.update(Klo, nk, i0, a - i0, b, Klo - b))
                        if a > i0 and b > j0:
                            new_rest = lambda args=(Klo, nk, i0, a - i0, j0, b - j0): self.update(*args)
                            rest_tiles += (a - i0) * (b - j0)
                    else:
                        stream.append(lambda: self.update(Klo, nk, i0, Klo - i0, j0, Klo - j0))
            if side is not None and side_first:
                rs, ws = side()
            R, W = set(), set()
            for f in stream:
                if isinstance(f, str):
                    if pending is not None:
                        join()
                    continue
                r, w = f()
                R |= r
                W |= w
                if pending is not None:
                    pending[1] |= r
                    pending[2] |= w
            if side is not None:
                if not side_first:
                    rs, ws = side()
                assert not (ws & (R | W)), "the second stream writes a tile the main stream uses"
                assert not (rs & W), "the second stream reads a tile the main stream writes"
            if new_rest is not None:  # started once this step's main-stream launches are queued
                pending = [new_rest, set(), set(), new_rest() if side_first else None]
        if pending is not None:
            join()
        return rest_tiles

    def check(self):
        N = self.NB * TB
        U = np.zeros((N, N), complex)
        L = np.eye(N, dtype=complex)
        for (I, J), t in self.T.items():
            s = np.s_[I * TB:(I + 1) * TB, J * TB:(J + 1) * TB]
            if I == J:
                U[s] = np.triu(t)
                L[s] = np.tril(t, -1) + np.eye(TB)
            elif I < J:
                U[s] = t
            else:
                L[s] = t
        return np.abs(U @ L - self.A).max() / np.abs(self.A).max()


@pytest.mark.parametrize("gk", [0, 1, 3, 5, 32])
@pytest.mark.parametrize("seed", [1, 2])
def test_schedule_reconstructs_a(gk, seed):
    A, lb1, lb2 = _matrix(330, seed)
    m = Model(A, lb1, lb2)
    m.factor(gk)
    assert m.check() <= 1e-12


@pytest.mark.parametrize("gk", [2, 3, 5, 32])
@pytest.mark.parametrize("seed", [3, 4])
def test_overlap_is_race_free_and_exact(gk, seed):
    A, lb1, lb2 = _matrix(330, seed)
    ref = Model(A, lb1, lb2)
    ref.factor(gk)
    for side_first in (False, True):
        m = Model(A, lb1, lb2)
        m.factor(gk, overlap=True, side_first=side_first)
        assert m.check() <= 1e-12
        for k in ref.T:  # the same factors up to the order of the accumulated terms
            assert np.abs(m.T[k] - ref.T[k]).max() <= 1e-12 * max(1.0, np.abs(ref.T[k]).max())


@pytest.mark.parametrize("gk", [3, 5])
@pytest.mark.parametrize("overlap", [False, True])
def test_lookahead_remainder_is_race_free(gk, overlap):
    """Wide profiles (bands of up to 20 tiles), so that groups leave a
    remainder beside the next group's chain: the same factors, and the
    remainder never touches a tile the main stream uses meanwhile."""
    A, lb1, lb2 = _matrix(560, 5, low_min=240, low_max=320, up_max=320, diag=400.0)
    ref = Model(A, lb1, lb2)
    ref.factor(gk)
    for side_first in (False, True):
        m = Model(A, lb1, lb2)
        assert m.factor(gk, overlap=overlap, side_first=side_first, lookahead=True) > 0
        assert m.check() <= 1e-12
        for k in ref.T:
            assert np.abs(m.T[k] - ref.T[k]).max() <= 1e-12 * max(1.0, np.abs(ref.T[k]).max())
