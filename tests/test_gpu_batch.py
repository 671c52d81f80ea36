"""Grouped device solver: several energy points and clusters of uneven size in
one grouped, graph-captured launch sequence; mixed precision; adjoint."""
import numpy as np
import pytest

from tests.helpers import dense_matrix, oracle_solution, small_system

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _run(precision, systems, repeats=1, max_sweeps=5, tol=1e-12, tune=None, overlap=None):
    import torch
    from fdmnes_amd.batch import DeviceSolver
    nrhs = systems[0]["rhs"].shape[1]
    cal = systems[0]["cal"]
    dt = np.complex128 if cal else np.float64
    sv = DeviceSolver(precision, cal, nrhs)
    if tune:
        sv.tune(**tune)
    if overlap is not None:
        sv.overlap(overlap)
    slots = [sv.add_system(s["lb1"], s["lb2"], s["rp"], s["col"]) for s in systems]
    vals = [_dev(s["vals"].astype(dt)) for s in systems]
    b = [_dev(s["rhs"].astype(dt)) for s in systems]
    x = [torch.zeros_like(t) for t in b]
    out = stages = sweeps = None
    if repeats > 1:
        sv.snapshots(True)
    for rep in range(repeats):  # graph replays must reproduce the first run stage by stage
        fr, sw, rel = sv.solve(slots, vals, b, x, max_sweeps=max_sweeps, tol=tol)
        cur = [t.cpu().numpy() for t in x]
        st = [sv.snapshot(k) for k in slots] if repeats > 1 else None
        if out is not None:
            assert sw == sweeps, (rep, sw, sweeps)
            for k, (p, q) in enumerate(zip(stages, st)):
                assert p["factor"] == q["factor"], f"replay {rep}: factor tiles of slot {k} differ"
                assert len(p["d"]) == len(q["d"])
                for j in range(len(p["d"])):
                    assert p["d"][j] == q["d"][j], f"replay {rep}: sweep {j} correction of slot {k} differs"
                    assert p["x"][j] == q["x"][j], f"replay {rep}: x after sweep {j} of slot {k} differs"
            for a, c in zip(out, cur):
                assert np.array_equal(a, c)
        out, stages, sweeps = cur, st, sw
    if repeats > 1:
        sv.snapshots(False)
    return sv, slots, out, sw, rel, b, x


def test_grouped_uneven_clusters_fp64():
    systems = [small_system(R=1.2, seed=1, E=0.3), small_system(R=1.6, seed=2, E=0.8),
               small_system(R=2.0, seed=3, E=1.5)]
    print("sizes", [len(s["lb1"]) for s in systems], flush=True)
    _, _, xs, _, _, _, _ = _run(0, systems, repeats=3)
    for s, x in zip(systems, xs):
        xo = oracle_solution(s)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_grouped_launch_of_24_uneven_systems(precision):
    """24 systems of six cluster sizes (R = 1.0 to 2.0 A) and four energies in
    one grouped solve, above the 16 of a half-full launch: every solution is
    the oracle's, and a replay of the captured graphs reproduces it."""
    radii = [1.0, 1.2, 1.4, 1.6, 1.8, 2.0]
    systems = [small_system(R=R, seed=10 + k, E=0.3 + 0.35 * (k % 4)) for k, R in enumerate(radii * 4)]
    print("24 systems, sizes", sorted({len(s["lb1"]) for s in systems}), flush=True)
    _, _, xs, sw, _, _, _ = _run(precision, systems, repeats=2)
    for s, x in zip(systems, xs):
        xo = oracle_solution(s)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


@pytest.mark.parametrize("lookahead", [0, 1, 96])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("trail_tile,gk,trail_wg", [(64, 0, 64), (64, 3, 64), (128, 3, 64), (128, 8, 64),
                                                     (128, 16, 64), (64, 2, 1 << 24), (64, 24, 1 << 24),
                                                     (64, 32, 1 << 24)])
def test_schedules_match_oracle(precision, trail_tile, gk, trail_wg, lookahead):
    """Every tile schedule (one-step updates, two-level blocking with 64 x 64
    or 2 x 2 macro output tiles, odd item edges, persistent grids or one
    workgroup per output block, groups up to the 32-step limit, with and
    without the lookahead split of each group's update over two streams)
    gives the oracle's solution; clusters of uneven size end their groups
    at different steps of the grouped launches."""
    systems = [small_system(R=1.2, seed=11, E=0.4), small_system(R=1.9, seed=12, E=1.1),
               small_system(R=2.2, seed=13, E=2.3)]
    print("precision", precision, "tile", trail_tile, "gk", gk, "wg", trail_wg, "lookahead", lookahead, "sizes",
          [len(s["lb1"]) for s in systems], flush=True)
    _, _, xs, sw, _, _, _ = _run(precision, systems, repeats=2,
                                  tune=dict(gk=gk, trail_wg=trail_wg, trail_tile=trail_tile, persist_solve=1,
                                            lookahead=lookahead))
    for s, x in zip(systems, xs):
        xo = oracle_solution(s)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("trail_tile,gk,trail_wg,lookahead", [(64, 3, 64, 0), (64, 8, 1 << 24, 0), (128, 8, 64, 0),
                                                              (64, 20, 1 << 24, 0), (64, 20, 1 << 24, 1),
                                                              (64, 32, 1 << 24, 96)])
def test_overlap_schedules_match_oracle(precision, trail_tile, gk, trail_wg, lookahead):
    """The per-step overlap (the next block row and column's left-looking
    update on the second stream beside the current panel), alone and with
    the group-level lookahead on the same stream, over groups of 3 to 32
    steps and clusters of uneven size: the oracle's solution, and graph
    replays that reproduce it."""
    systems = [small_system(R=1.2, seed=11, E=0.4), small_system(R=1.9, seed=12, E=1.1),
               small_system(R=2.2, seed=13, E=2.3)]
    print("overlap: precision", precision, "tile", trail_tile, "gk", gk, "wg", trail_wg, "lookahead", lookahead,
          flush=True)
    _, _, xs, _, _, _, _ = _run(precision, systems, repeats=3, overlap=True,
                                tune=dict(gk=gk, trail_wg=trail_wg, trail_tile=trail_tile, persist_solve=1,
                                          lookahead=lookahead))
    for s, x in zip(systems, xs):
        xo = oracle_solution(s)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


@pytest.mark.parametrize("trail_tile", [64, 128])
def test_real_arithmetic_schedules(trail_tile):
    """Real arithmetic (Basereel without complex energy) through both update kernels."""
    systems = [small_system(R=1.6, seed=14, E=0.9, E_comp=False, basereel=True),
               small_system(R=2.0, seed=15, E=1.7, E_comp=False, basereel=True)]
    assert not systems[0]["cal"]
    for precision in (0, 1):
        print("real precision", precision, "tile", trail_tile, flush=True)
        _, _, xs, _, _, _, _ = _run(precision, systems, repeats=2, tune=dict(gk=4, trail_wg=64, trail_tile=trail_tile))
        for s, x in zip(systems, xs):
            xo = oracle_solution(s)
            assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


def test_grouped_mixed_precision():
    systems = [small_system(R=1.6, seed=4, E=e) for e in (0.2, 0.9, 2.0, 3.5)]
    _, _, xs, sw, rel, _, _ = _run(1, systems, repeats=2)
    print("sweeps", sw, "rel", rel, flush=True)
    for s, x, k in zip(systems, xs, sw):
        assert k >= 0
        xo = oracle_solution(s)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10


def test_sigma_matches_numpy():
    import torch
    systems = [small_system(R=1.6, seed=5, E=0.7)]
    sv, slots, xs, _, _, b, x = _run(0, systems)
    sig = sv.sigma(slots, b, x)
    ref = -np.imag(np.sum(systems[0]["rhs"] * xs[0]))
    assert abs(sig[0] - ref) <= 1e-12 * abs(ref)


def test_adjoint_vs_dense_gradient_and_fd():
    import torch
    s = small_system(R=1.4, seed=6, E=0.7, lmaxso=2)
    sv, slots, xs, _, _, b, x = _run(0, [s])
    n = len(s["lb1"])
    grad = torch.zeros(n, dtype=torch.float64, device="cuda")
    sv.adjoint(slots, b, x, [grad])
    g = grad.cpu().numpy()
    # FP64 CPU dense-solve gradient: A x = s, A^T lam = s, dsigma/dV_i = Im sum_L lam_iL x_iL
    A = dense_matrix(s)
    xs_d = np.linalg.solve(A, s["rhs"])
    lam = np.linalg.solve(A.T, s["rhs"])
    g_ref = np.imag(np.sum(lam * xs_d, axis=1))
    npt = s["cluster"].npoint
    assert np.abs(g[:npt] - g_ref[:npt]).max() <= 1e-9 * np.abs(g_ref).max()
    # central finite differences through the GPU forward solve on random grid points
    from fdmnes_amd import configs as cfg
    rng = np.random.default_rng(0)
    dV = 1e-6
    cl = s["cluster"]
    for i in rng.choice(npt, 8, replace=False):
        sig = []
        for sgn in (1, -1):
            V = s["V"].copy()
            V[i] += sgn * dV
            rp, col, vals, _ = cl.assemble(V, s["E"], s["eimag"], cfg.ecinetic_out(s["E"]), True, False)
            t = small_system.__defaults__  # noqa: F841
            vv = _dev(vals.astype(np.complex128))
            xx = torch.zeros_like(b[0])
            sv.solve(slots, [vv], b, [xx])
            sig.append(sv.sigma(slots, b, [xx])[0])
        fd = (sig[0] - sig[1]) / (2 * dV)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max() + 1e-7 * abs(g[i]), (i, fd, g[i])
