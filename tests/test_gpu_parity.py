"""GPU FP64 drop-in vs the oracle's mat_solve restatement on the same inputs."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

BOHR = 0.52917721067


def _case(R, h, lmaxso, E_comp, cal, basereel, seed=0, E=0.6, eimag=0.0367):
    cl = O.Cluster(R / BOHR, h / BOHR, lmaxso)
    rng = np.random.default_rng(seed)
    V = rng.uniform(-0.3, 0.3, cl.npoint) - 0.8
    Eimag = eimag if E_comp else 0.0
    rp, col, vals, b_inc = cl.assemble_csr(V, E, Eimag, E, E_comp, cal, basereel)
    lb1, lb2, _, _ = cl.profile()
    rhs = b_inc
    st, x_or, _, _, _, _ = cl.solve(V, E, Eimag, E, E_comp, cal, basereel)
    assert st == 0
    return cl, lb1, lb2, rp, col, vals, rhs, x_or


@pytest.mark.parametrize("R,h,lm,E_comp,cal,basereel", [
    (1.2, 0.25, 2, True, True, False),
    (1.2, 0.25, 2, False, False, True),
    (1.2, 0.25, 2, False, True, True),
    (2.0, 0.25, 3, True, True, False),
    (2.6, 0.3, 3, False, False, True),
])
def test_fp64_matches_oracle(R, h, lm, E_comp, cal, basereel):
    from fdmnes_amd import solver
    cl, lb1, lb2, rp, col, vals, rhs, x_or = _case(R, h, lm, E_comp, cal, basereel)
    print(f"case R={R} h={h} n={cl.nligne} E_comp={E_comp} cal={cal}", flush=True)
    x, tf, tt = solver.mat_solve(lb1, lb2, rp, col, vals, rhs, cal)
    rel = np.abs(x - x_or).max() / np.abs(x_or).max()
    assert rel <= 1e-10, rel


def test_mixed_matches_oracle():
    from fdmnes_amd import solver
    cl, lb1, lb2, rp, col, vals, rhs, x_or = _case(2.0, 0.25, 3, True, True, False)
    x, sweeps, rel_res = solver.mat_solve_mixed(lb1, lb2, rp, col, vals, rhs, True)
    print("sweeps", sweeps, "rel", rel_res, flush=True)
    rel = np.abs(x - x_or).max() / np.abs(x_or).max()
    assert sweeps >= 0 and rel_res <= 1e-12
    assert rel <= 1e-10, rel


def test_dropin_reuses_its_solver_across_energies():
    """The drop-in keeps its device solver while the structure stays the same
    (FDMNES's energy loop): consecutive energies of one cluster, another
    cluster in between, and the first energy again all give the oracle's
    solutions, and a repeated call gives the same bits."""
    from fdmnes_amd import solver
    cases = [_case(2.0, 0.25, 3, True, True, False, E=e) for e in (0.4, 1.1)]
    other = _case(1.6, 0.25, 2, True, True, False, E=0.8)
    seq = [cases[0], cases[1], other, cases[0], cases[1]]
    first = None
    for k, (cl, lb1, lb2, rp, col, vals, rhs, x_or) in enumerate(seq):
        print(f"call {k}: n={cl.nligne}", flush=True)
        x, _, _ = solver.mat_solve(lb1, lb2, rp, col, vals, rhs, True)
        rel = np.abs(x - x_or).max() / np.abs(x_or).max()
        assert rel <= 1e-10, (k, rel)
        if k == 0:
            first = x.copy()
        if k == 3:
            assert np.array_equal(x, first)
    # mixed precision over the same sequence (its own cached solver)
    for k, (cl, lb1, lb2, rp, col, vals, rhs, x_or) in enumerate(seq[:3]):
        x, sweeps, rel_res = solver.mat_solve_mixed(lb1, lb2, rp, col, vals, rhs, True)
        assert rel_res <= 1e-12 and np.abs(x - x_or).max() / np.abs(x_or).max() <= 1e-10, k
