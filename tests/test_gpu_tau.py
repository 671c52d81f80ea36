"""Device post-solve reduction (fdmgpu_tau_reduce) against the oracle's
restatement of mat.f90:297-405 on the same solutions, its chunking and error
path, and end to end: solutions of the real (Neumann) outer-sphere basis,
renormalised with invcomp, give the tau of the complex (Bessel + i Neumann)
basis (the identity the oracle's tests establish), for the FP64 and the mixed
solver, on a small cluster and on the full-size copper cluster."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
BOHR = 0.52917721067


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _oracle_tau(x, nso, basereel, fdm_comp, atoms):
    t, norm, info = O.tau(x, nso, basereel, fdm_comp, atoms, want_norm=basereel)
    assert info == 0
    return np.concatenate([b.reshape(-1) for b in t]), norm


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nl", [9, 16, 64])
@pytest.mark.parametrize("cal", [False, True])
@pytest.mark.parametrize("basereel,fdm_comp", [(True, False), (True, True), (False, False), (False, True)])
def test_tau_matches_oracle(nl, cal, basereel, fdm_comp):
    from fdmnes_amd.tau import tau_reduce
    print(f"tau nlmso={nl} cal={cal} basereel={basereel} fdm_comp={fdm_comp}", flush=True)
    rng = np.random.default_rng(nl + 2 * cal + 4 * basereel + 8 * fdm_comp)
    n = 3 * nl + 40
    xs = [rng.normal(size=(n, nl)) + (1j * rng.normal(size=(n, nl)) if cal else 0) for _ in range(3)]
    atoms = [(0, min(nl, 16)), (7, 4), (n - nl - 3, 3), (n - 1, 1)]
    tau, norm, info = tau_reduce([_dev(x) for x in xs], n - nl, basereel, fdm_comp, atoms, want_norm=True)
    assert info.cpu().tolist() == [0, 0, 0]
    exact = True
    for k, x in enumerate(xs):
        ot, on = _oracle_tau(x, n - nl, basereel, fdm_comp, atoms)
        dt = tau[k].cpu().numpy()
        assert _rel(dt, ot) <= 1e-13
        exact &= np.array_equal(dt, ot)
        if basereel:
            dn = norm[k].cpu().numpy().T  # column-major on the device
            assert _rel(dn, on) <= 1e-13
            exact &= np.array_equal(dn, on)
    print("bit-exact vs oracle:", exact, flush=True)


def test_tau_accumulates_and_chunks():
    """20 points (two launches of up to 16) and 40 atoms (two atom chunks of
    up to 32), accumulated onto earlier contents."""
    import torch
    from fdmnes_amd.tau import tau_reduce
    rng = np.random.default_rng(11)
    n, nl = 120, 16
    xs = [rng.normal(size=(n, nl)) + 1j * rng.normal(size=(n, nl)) for _ in range(20)]
    atoms = [(int(r), int(k)) for r, k in zip(rng.integers(0, n - 9, 40), rng.integers(1, 9, 40))]
    tot = sum(k * k for _, k in atoms)
    prior = [rng.normal(size=tot) + 1j * rng.normal(size=tot) for _ in xs]
    buf = [_dev(p) for p in prior]
    tau, _, info = tau_reduce([_dev(x) for x in xs], n - nl, True, False, atoms, tau=buf, accumulate=True)
    assert int(info.abs().sum()) == 0
    for k, x in enumerate(xs):
        ot, _ = _oracle_tau(x, n - nl, True, False, atoms)
        assert _rel(tau[k].cpu().numpy(), prior[k] + ot) <= 1e-13
    # a third point set without accumulation overwrites
    tau2, _, _ = tau_reduce([_dev(x) for x in xs[:2]], n - nl, True, False, atoms, tau=buf[:2])
    torch.cuda.synchronize()
    ot, _ = _oracle_tau(xs[0], n - nl, True, False, atoms)
    assert _rel(tau2[0].cpu().numpy(), ot) <= 1e-13


def test_singular_normalisation_is_reported():
    """norm = I + i X with X = diag(1, i, 1, 1) is singular at pivot 2: the
    status names it, the point's tau is left alone, the others are reduced."""
    from fdmnes_amd import _lib
    from fdmnes_amd.tau import tau_reduce
    n, nl = 20, 4
    good = np.random.default_rng(1).normal(size=(n, nl)) + 0j
    bad = np.zeros((n, nl), complex)
    bad[n - nl:, :] = np.diag([1, 1j, 1, 1])
    atoms = [(0, 3)]
    tau, _, info = tau_reduce([_dev(good), _dev(bad)], n - nl, True, False, atoms, check=False)
    assert info.cpu().tolist() == [0, 2]
    assert np.all(tau[1].cpu().numpy() == 0)
    ot, _ = _oracle_tau(good, n - nl, True, False, atoms)
    assert _rel(tau[0].cpu().numpy(), ot) <= 1e-13
    with pytest.raises(_lib.FdmGpuError):
        tau_reduce([_dev(good), _dev(bad)], n - nl, True, False, atoms)


def test_tau_rejects_bad_arguments():
    from fdmnes_amd import _lib
    from fdmnes_amd.tau import tau_reduce
    x = _dev(np.zeros((30, 9)))
    with pytest.raises(_lib.FdmGpuError):  # atom rows beyond the system
        tau_reduce([x], 21, True, False, [(25, 9)])
    with pytest.raises(_lib.FdmGpuError):  # outer rows beyond the system
        tau_reduce([x], 25, True, False, [(0, 3)])


def _basis_pair(cl, V, E, Ec, prec):
    """Device solutions of one real energy in the real and in the complex
    outer-sphere basis (host-assembled rows, incoming-wave RHS)."""
    import torch
    from fdmnes_amd.batch import DeviceSolver
    lb1, lb2 = cl.profile()
    out = []
    for basereel in (True, False):
        rp, col, vals, rhs = cl.assemble(V, E, 0.0, Ec, False, basereel)
        cal = not basereel
        sv = DeviceSolver(prec, cal, cl.nlmso)
        slot = sv.add_system(lb1, lb2, rp, col)
        b = _dev(rhs)
        x = torch.zeros_like(b)
        sv.solve([slot], [_dev(vals)], [b], [x])
        torch.cuda.synchronize()
        out.append(x)
        del sv
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_real_basis_tau_equals_complex_basis_tau(prec):
    """Bound 1e-9: five times what the reference algorithm itself reaches on
    this case (its unpivoted elimination of the real-basis system, cond
    5e3, leaves the identity at 1.6e-10 and 1.9e-10 with and without
    FDM_comp, measured with the oracle, which equals the reference bit for
    bit on the real branches; pivoted dense solves give 4e-14)."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.tau import tau_reduce
    print(f"basis identity, small cluster, precision {prec}", flush=True)
    cl = CL.Cluster(2.0 / BOHR, 0.25 / BOHR, 3)
    V = np.random.default_rng(7).uniform(-1.5, 0.0, cl.npoint)
    xr, xc = _basis_pair(cl, V, 0.45, 0.7, prec)
    nso = cl.nligne - cl.nlmso
    atoms = [(0, 16), (100, 9), (nso - 4, 4)]
    for fdm_comp in (False, True):
        tr, _, _ = tau_reduce([xr], nso, True, fdm_comp, atoms)
        tc, _, _ = tau_reduce([xc], nso, False, fdm_comp, atoms)
        a, b = tr[0].cpu().numpy(), tc[0].cpu().numpy()
        print("fdm_comp", fdm_comp, "real vs complex basis", _rel(a, b), flush=True)
        assert _rel(a, b) <= 1e-9
        # and the device reduction equals the oracle's on the same solutions
        ot, _ = _oracle_tau(xr.cpu().numpy(), nso, True, fdm_comp, atoms)
        assert _rel(a, ot) <= 1e-13


def test_full_size_copper_basis_identity():
    """The 91-energy copper workload's cluster (R = 3 A, h = 0.24 A, l <= 2,
    8,226 unknowns) at its middle energy: tau from the real basis with
    renormalisation equals tau from the complex basis.  Bound 2e-10: five
    times what the reference algorithm reaches here (3.8e-11, the oracle's
    band solves of both bases)."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    from fdmnes_amd.tau import tau_reduce
    w = cfg.CU_REAL
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    V = cfg.potential(w, cl.points)
    E = float(w.energies_Ry[len(w.energies_Ry) // 2])
    print("copper basis identity nligne", cl.nligne, flush=True)
    xr, xc = _basis_pair(cl, V, E, cfg.ecinetic_out(E), 0)
    nso = cl.nligne - cl.nlmso
    atoms = [(0, 9), (nso // 2, 9), (nso - 9, 9)]
    tr, _, _ = tau_reduce([xr], nso, True, False, atoms)
    tc, _, _ = tau_reduce([xc], nso, False, False, atoms)
    a, b = tr[0].cpu().numpy(), tc[0].cpu().numpy()
    print("real vs complex basis", _rel(a, b), flush=True)
    assert _rel(a, b) <= 2e-10
    ot, _ = _oracle_tau(xr.cpu().numpy(), nso, True, False, atoms)
    assert _rel(a, ot) <= 1e-13


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_device_pipeline_matches_oracle_chain(prec, E_comp, basereel):
    """The whole per-energy path on the GPU (device assembly, grouped solve,
    tau reduction) against the oracle's chain (calcMatRow and mat_solve
    restatements, then the reduction) for three energies of a small
    cluster.  Bounds: 1e-10 in the complex basis (the solver's parity bar);
    1e-9 in the real basis, where the reduction inverts I + i X (its
    unpivoted solve leaves 2e-10 against a pivoted one on such clusters)."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    from fdmnes_amd.batch import DeviceSolver
    from fdmnes_amd.tau import tau_reduce
    print(f"pipeline precision {prec} E_comp {E_comp} basereel {basereel}", flush=True)
    R, h, lm = 1.6 / BOHR, 0.25 / BOHR, 2
    c, o = CL.Cluster(R, h, lm), O.Cluster(R, h, lm)
    cal = E_comp or not basereel
    rng = np.random.default_rng(21)
    ens = [(-0.2 + 0.35 * k, 0.0367 if E_comp else 0.0, 0.3 + 0.4 * k) for k in range(3)]
    V = [rng.uniform(-1.5, 0.0, c.npoint) for _ in ens]
    dt = torch.complex128 if cal else torch.float64
    vals = [torch.zeros(c.nnz, dtype=dt, device="cuda") for _ in ens]
    rhs = [torch.zeros((c.nligne, c.nlmso), dtype=dt, device="cuda") for _ in ens]
    DeviceAssembler(c).run([_dev(v) for v in V], ens, E_comp, basereel, vals, rhs)
    lb1, lb2 = c.profile()
    rp, col, _, _ = c.assemble(V[0], *ens[0], E_comp, basereel)
    sv = DeviceSolver(prec, cal, c.nlmso)
    slots = [sv.add_system(lb1, lb2, rp, col) for _ in ens]
    xs = [torch.zeros_like(b) for b in rhs]
    sv.solve(slots, vals, rhs, xs)
    nso = c.nligne - c.nlmso
    atoms = [(0, 9), (200, 4), (nso - 3, 3)]
    tau, _, info = tau_reduce(xs, nso, basereel, False, atoms)
    bound = 1e-9 if basereel else 1e-10
    for k, (E, Ei, Ec) in enumerate(ens):
        st, xo, *_ = o.solve(V[k], E, Ei, Ec, E_comp, cal, basereel)
        assert st == 0
        ot, _ = _oracle_tau(xo, nso, basereel, False, atoms)
        d = _rel(tau[k].cpu().numpy(), ot)
        print("point", k, "tau vs oracle chain", d, flush=True)
        assert d <= bound
