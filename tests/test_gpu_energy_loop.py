"""The device energy loop (fdmnes_amd.energy_loop.EnergyLoop) against the
oracle's chain (calcMatRow and mat_solve restatements with calcMatRow's
incoming-wave RHS, then the post-solve reduction) energy by energy: five
energies in groups of two (a partial last group; the rows of the later
groups updated in place), both arithmetic branches, both precisions.
Bounds: solutions and tau 1e-10 in the complex basis (the solver's parity
bar); 1e-9 for tau in the real basis, where the reduction inverts I + i X."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
BOHR = 0.52917721067


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_energy_loop_matches_oracle_chain(precision, E_comp, basereel):
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import EnergyLoop
    print(f"energy loop precision {precision} E_comp {E_comp} basereel {basereel}", flush=True)
    R, h, lm = 1.6 / BOHR, 0.25 / BOHR, 2
    c, o = CL.Cluster(R, h, lm), O.Cluster(R, h, lm)
    cal = E_comp or not basereel
    rng = np.random.default_rng(31)
    ens = [(-0.25 + 0.3 * k, 0.0367 if E_comp else 0.0, 0.25 + 0.35 * k) for k in range(5)]
    V = rng.uniform(-1.5, 0.0, c.npoint)
    nso = c.nligne - c.nlmso
    atoms = [(0, 9), (150, 4)]
    loop = EnergyLoop(c, precision=precision, group=2, E_comp=E_comp, basereel=basereel, atoms=atoms)
    out = loop.run(V, ens, keep_solutions=True)
    assert len(out["tau"]) == len(ens) and len(out["x"]) == len(ens)
    for k, (E, Ei, Ec) in enumerate(ens):
        st, xo, *_ = o.solve(V, E, Ei, Ec, E_comp, cal, basereel)
        assert st == 0
        _, _, _, b = o.assemble_csr(V, E, Ei, Ec, E_comp, cal, basereel)
        dx = _rel(out["x"][k], xo)
        sig_o = -np.imag(np.sum(b * xo))
        to, _, info = O.tau(xo, nso, basereel, False, atoms)
        assert info == 0
        dt = max(_rel(a, b_) for a, b_ in zip(out["tau"][k], to))
        print("energy", k, "x", dx, "tau", dt, "sigma", out["sigma"][k], sig_o, flush=True)
        assert dx <= (1e-9 if basereel else 1e-10)
        assert dt <= (1e-9 if basereel else 1e-10)
        assert abs(out["sigma"][k] - sig_o) <= 1e-10 * max(abs(sig_o), np.abs(b).max() * np.abs(xo).max())


def test_energy_loop_potential_per_energy_and_reuse():
    """One potential per energy, and a second run on the same loop object
    (its buffers already hold assembled rows) gives the same as a fresh one."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import EnergyLoop
    c = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    rng = np.random.default_rng(32)
    ens = [(0.1 + 0.2 * k, 0.0367, 0.4 + 0.2 * k) for k in range(3)]
    Vs = [rng.uniform(-1.5, 0.0, c.npoint) for _ in ens]
    atoms = [(0, 9)]
    loop = EnergyLoop(c, precision=0, group=2, atoms=atoms)
    a = loop.run(Vs, ens)
    b = loop.run(Vs[::-1], ens[::-1])
    fresh = EnergyLoop(c, precision=0, group=2, atoms=atoms).run(Vs[::-1], ens[::-1])
    assert np.allclose(b["sigma"], fresh["sigma"], rtol=1e-13, atol=0)
    for t1, t2 in zip(b["tau"], fresh["tau"]):
        assert _rel(t1[0], t2[0]) <= 1e-13
    assert np.allclose(a["sigma"], b["sigma"][::-1], rtol=1e-12, atol=0)


def test_energy_loop_gradient():
    """dsigma/dV from the FP64 loop's adjoint against central differences of
    sigma through the same loop (grid points near the centre, the boundary
    and in between) and against a dense CPU adjoint on the oracle's rows."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import FP64, EnergyLoop
    c = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    o = O.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    rng = np.random.default_rng(41)
    ens = [(0.1 + 0.3 * k, 0.0367, 0.4 + 0.3 * k) for k in range(3)]
    V = rng.uniform(-1.5, 0.0, c.npoint)
    loop = EnergyLoop(c, precision=FP64, group=2)
    out = loop.run(V, ens, gradient=True)
    g = out["dsigma_dV"]
    assert g.shape == (3, c.npoint)
    h = 1e-4
    for i in (0, c.npoint // 2, c.npoint - 1):
        vp, vm = V.copy(), V.copy()
        vp[i] += h
        vm[i] -= h
        fd = (loop.run(vp, ens)["sigma"] - loop.run(vm, ens)["sigma"]) / (2 * h)
        print("point", i, "adjoint", g[:, i], "central differences", fd, flush=True)
        assert np.abs(g[:, i] - fd).max() <= 1e-6 * np.abs(g).max()
    # dense CPU adjoint of the oracle's rows: A^T lam = b, dsigma/dV_i = Im sum_L lam_iL x_iL
    for k, e in enumerate(ens):
        A, b = o.assemble_dense(V, *e, True, True, False)
        x = np.linalg.solve(A, b)
        lam = np.linalg.solve(A.T, b)
        ref = np.imag(np.sum(lam * x, axis=1))[:c.npoint]
        assert np.abs(g[k] - ref).max() <= 1e-9 * np.abs(ref).max()
