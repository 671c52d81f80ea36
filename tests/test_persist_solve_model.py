"""A model of the persistent block substitution (k_solve_persist) on the CPU:
its ticket order, the dependency order of each task (older blocks first, the
neighbour last), the done-flag protocol and the XCD-local L2 caches, run under
random interleavings of workgroups.

Memory: each XCD has an L2 that keeps a workgroup's stores of x_K (dirty)
until a write-back copies every dirty block of that L2 to memory; a consumer
reads a block from its own XCD's L2 when it holds it dirty, else from memory.
A task K publishes x_K as the kernel does: store x_K, write the L2 back, wait
for the write-back, then store done[K].  A consumer waits on done[J] (the
acquire invalidates its clean copies) before it reads x_J.

Checks, for the four substitution kinds over clusters of uneven size aligned
at their last block step (as the grouped launches align them):
- every interleaving gives numpy's block triangular solution, with one to six
  workgroups on three XCDs, and no task ever waits on a larger ticket;
- a flag stored before the write-back has completed (the flag's release
  without its wait), or a neighbour block read before its flag, gives a wrong
  solution under some interleaving: the model can tell."""
import numpy as np
import pytest

TB = 4
NXCD = 3
S_LOWER, S_UPPER, S_LOWER_T, S_UPPER_T = 0, 1, 2, 3


def _plan(rng, NB):
    """Block skyline with J0 non-decreasing and J1 swept non-decreasing."""
    low = rng.integers(0, 4, NB)
    J0 = np.maximum.accumulate(np.maximum(0, np.arange(NB) - low))
    J1 = np.maximum.accumulate(np.minimum(NB - 1, np.arange(NB) + rng.integers(0, 4, NB)))
    I0 = np.array([int(np.argmax(J1 >= K)) for K in range(NB)])
    Jx = np.array([int(np.nonzero(J0 <= K)[0].max()) for K in range(NB)])
    return J0, J1, I0, Jx


def _range(kind, K, J0, J1, I0, Jx):
    if kind == S_LOWER:
        return J0[K], K
    if kind == S_UPPER:
        return K + 1, J1[K] + 1
    if kind == S_LOWER_T:
        return K + 1, Jx[K] + 1
    return I0[K], K


class System:
    def __init__(self, rng, NB):
        self.NB = NB
        self.J0, self.J1, self.I0, self.Jx = _plan(rng, NB)
        n = NB * TB
        A = np.zeros((n, n))
        for I in range(NB):
            for J in range(self.J0[I], self.J1[I] + 1):
                A[I * TB:(I + 1) * TB, J * TB:(J + 1) * TB] = rng.normal(size=(TB, TB))
        A += 4 * TB * np.eye(n)
        self.A = A
        self.b = rng.normal(size=(n, 2))

    def block(self, I, J):
        return self.A[I * TB:(I + 1) * TB, J * TB:(J + 1) * TB]

    def operator(self, kind):
        """The triangular matrix the substitution of this kind solves with."""
        n = self.NB * TB
        M = np.zeros((n, n))
        for K in range(self.NB):
            jb, je = _range(kind, K, self.J0, self.J1, self.I0, self.Jx)
            rows = slice(K * TB, (K + 1) * TB)
            for J in list(range(jb, je)) + [K]:
                T = self.block(J, K).T if kind in (S_LOWER_T, S_UPPER_T) else self.block(K, J)
                M[rows, J * TB:(J + 1) * TB] = T
        return M


def _deps(kind, K, s):
    """Older dependencies in the order the kernel accumulates them, then the
    neighbour (None without one)."""
    jb, je = _range(kind, K, s.J0, s.J1, s.I0, s.Jx)
    if je <= jb:
        return [], None
    asc = kind in (S_LOWER, S_UPPER_T)
    if asc:
        return list(range(jb, je - 1)), je - 1
    return list(range(je - 1, jb, -1)), jb


def _tasks(systems, kind):
    """(system, K) by ticket: step by step, systems in list order within a
    step, every system ending at the last step."""
    maxNB = max(s.NB for s in systems)
    asc = kind in (S_LOWER, S_UPPER_T)
    out = []
    for step in range(maxNB):
        for i, s in enumerate(systems):
            t = step - (maxNB - s.NB)
            if 0 <= t < s.NB:
                out.append((i, t if asc else s.NB - 1 - t))
    return out


def run_model(systems, kind, nwg, rng, flag_first=False, early_neighbour=False):
    """One random interleaving of nwg workgroups; returns the solution per
    system (x blocks as memory holds them at the end)."""
    trans = kind in (S_LOWER_T, S_UPPER_T)
    tasks = _tasks(systems, kind)
    ticket_of = {t: k for k, t in enumerate(tasks)}
    memory = {(i, K): s.b[K * TB:(K + 1) * TB].copy() for i, s in enumerate(systems) for K in range(s.NB)}
    flags = {key: False for key in memory}
    l2 = [dict() for _ in range(NXCD)]  # dirty blocks per XCD
    counter = [0]

    def read(xcd, key):
        return (l2[xcd][key] if key in l2[xcd] else memory[key]).copy()

    def writeback(xcd):
        memory.update(l2[xcd])
        l2[xcd].clear()

    def workgroup(w):
        xcd = w % NXCD
        while True:
            t = counter[0]
            counter[0] += 1
            yield
            if t >= len(tasks):
                return
            i, K = tasks[t]
            s = systems[i]
            older, nb = _deps(kind, K, s)

            def wait(J):
                assert ticket_of[(i, J)] < t, "a task waits only on smaller tickets"
                while not flags[(i, J)]:
                    yield
            T = (lambda J: s.block(J, K).T) if trans else (lambda J: s.block(K, J))
            acc = np.zeros((TB, 2))
            for J in older:
                yield from wait(J)
                acc += T(J) @ read(xcd, (i, J))
                yield
            r = read(xcd, (i, K))
            if nb is not None:
                if early_neighbour:
                    xnb = read(xcd, (i, nb))
                    yield
                    yield from wait(nb)
                else:
                    yield from wait(nb)
                    xnb = read(xcd, (i, nb))
                acc += T(nb) @ xnb
            x = np.linalg.solve(s.block(K, K).T if trans else s.block(K, K), r - acc)
            l2[xcd][(i, K)] = x
            yield
            if flag_first:  # the flag leaves before the write-back completes
                flags[(i, K)] = True
                yield
                writeback(xcd)
            else:
                writeback(xcd)
                yield
                flags[(i, K)] = True
            yield

    live = [workgroup(w) for w in range(nwg)]
    for _ in range(200000):
        if not live:
            break
        k = int(rng.integers(len(live)))
        try:
            next(live[k])
        except StopIteration:
            live.pop(k)
    assert not live, "the persistent solve did not drain"
    for x in l2:
        memory.update(x)  # kernel end: every L2 written back
    return [np.concatenate([memory[(i, K)] for K in range(s.NB)]) for i, s in enumerate(systems)]


def _systems(seed):
    rng = np.random.default_rng(seed)
    return [System(rng, NB) for NB in (7, 4, 9)]


@pytest.mark.parametrize("kind", [S_LOWER, S_UPPER, S_LOWER_T, S_UPPER_T])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_persistent_solve_matches_numpy_under_any_interleaving(kind, seed):
    systems = _systems(seed)
    rng = np.random.default_rng(100 + seed)
    refs = [np.linalg.solve(s.operator(kind), s.b) for s in systems]
    for trial in range(40):
        xs = run_model(systems, kind, int(rng.integers(1, 7)), rng)
        for x, ref in zip(xs, refs):
            assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max(), (kind, seed, trial)


def _some_interleaving_fails(**planted):
    for seed in range(6):
        systems = _systems(seed)
        rng = np.random.default_rng(seed)
        for kind in (S_LOWER, S_UPPER):
            refs = [np.linalg.solve(s.operator(kind), s.b) for s in systems]
            for _ in range(30):
                xs = run_model(systems, kind, int(rng.integers(2, 7)), rng, **planted)
                if any(np.abs(x - r).max() > 1e-10 * np.abs(r).max() for x, r in zip(xs, refs)):
                    return True
    return False


def test_flag_before_write_back_is_caught():
    assert _some_interleaving_fails(flag_first=True)


def test_neighbour_read_before_its_flag_is_caught():
    assert _some_interleaving_fails(early_neighbour=True)
