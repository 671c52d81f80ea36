"""Device-side per-energy assembly (fdmgpu_assembler_*): the values and RHS it
writes against the host assembler and the oracle's calcMatRow rows, for the
three arithmetic branches (mat.f90:96-104), and the solver run on them."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
BOHR = 0.52917721067

MODES = [(True, False), (False, True), (False, False), (True, True)]  # (E_comp, basereel)


def _energies(n, E_comp):
    # (Enervide, Eimag, Ecinetic_out) in Ry, spread over the workloads' range
    return [(-0.3 + 0.4 * k, 0.0367 if E_comp else 0.0, 0.2 + 0.5 * k) for k in range(n)]


def _device_run(c, V, ens, E_comp, basereel, with_rhs=True):
    import torch
    from fdmnes_amd.assemble import DeviceAssembler
    cal = E_comp or not basereel
    dt = torch.complex128 if cal else torch.float64
    a = DeviceAssembler(c)
    Vd = [torch.from_numpy(v).cuda() for v in V]
    vals = [torch.full((c.nnz,), 7.0, dtype=dt, device="cuda") for _ in ens]  # stale contents to overwrite
    rhs = [torch.full((c.nligne, c.nlmso), 7.0, dtype=dt, device="cuda") for _ in ens] if with_rhs else None
    a.run(Vd, ens, E_comp, basereel, vals, rhs)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in vals], ([t.cpu().numpy() for t in rhs] if with_rhs else None)


def _fixed_mask(c, t):
    m = np.ones(c.nnz, bool)
    m[t["qdiag"]] = False
    for r in range(t["nrow_out"]):
        m[t["oq"][r]: t["oq"][r] + c.nlmso] = False
    for lm in range(c.nlmso):
        q = t["row_ptr"][c.npoint + lm] + c.nsortf
        m[q: q + c.nlmso] = False
    return m


@pytest.mark.parametrize("R,h,lm", [(1.6, 0.25, 2), (2.4, 0.2, 3)])
@pytest.mark.parametrize("E_comp,basereel", MODES)
def test_device_assembly_matches_host_and_oracle(R, h, lm, E_comp, basereel):
    from fdmnes_amd import cluster as CL
    print(f"assembly R={R} h={h} lmax={lm} E_comp={E_comp} basereel={basereel}", flush=True)
    c = CL.Cluster(R / BOHR, h / BOHR, lm)
    o = O.Cluster(R / BOHR, h / BOHR, lm)
    t = c.assembly_tables()
    rng = np.random.default_rng(5)
    ens = _energies(3, E_comp)
    V = [rng.uniform(-1.5, 0.0, c.npoint) for _ in ens]
    dv, dr = _device_run(c, V, ens, E_comp, basereel)
    cal = E_comp or not basereel
    fixed = _fixed_mask(c, t)
    for k, (E, Ei, Ec) in enumerate(ens):
        _, _, hv, hr = c.assemble(V[k], E, Ei, Ec, E_comp, basereel)
        _, _, ov, orr = o.assemble_csr(V[k], E, Ei, Ec, E_comp, cal, basereel)
        assert dv[k].dtype == (np.complex128 if cal else np.float64)
        # energy-independent entries and grid diagonals: bit for bit
        assert np.array_equal(dv[k][fixed], hv[fixed])
        assert np.array_equal(dv[k][t["qdiag"]], hv[t["qdiag"]])
        # entries made of outer-sphere functions: to the last bits of j_l, n_l
        scale = np.abs(hv).max()
        assert np.abs(dv[k] - hv).max() <= 1e-13 * scale
        assert np.abs(dv[k] - ov).max() <= 1e-13 * np.abs(ov).max()
        assert np.abs(dr[k] - hr).max() <= 1e-13 * np.abs(hr).max()
        assert np.abs(dr[k] - orr).max() <= 1e-13 * np.abs(orr).max()


@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_update_only_rewrites_every_energy_entry(E_comp, basereel):
    """A second call on the same buffers with update_only (new energies and
    potentials) gives exactly what a call on fresh buffers gives."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    c = CL.Cluster(2.4 / BOHR, 0.2 / BOHR, 3)
    cal = E_comp or not basereel
    dt = torch.complex128 if cal else torch.float64
    rng = np.random.default_rng(8)
    a = DeviceAssembler(c)
    e1, e2 = _energies(3, E_comp), _energies(6, E_comp)[3:]
    V1 = [torch.from_numpy(rng.uniform(-1.5, 0.0, c.npoint)).cuda() for _ in e1]
    V2 = [torch.from_numpy(rng.uniform(-1.5, 0.0, c.npoint)).cuda() for _ in e2]
    vals = [torch.zeros(c.nnz, dtype=dt, device="cuda") for _ in e1]
    rhs = [torch.zeros((c.nligne, c.nlmso), dtype=dt, device="cuda") for _ in e1]
    a.run(V1, e1, E_comp, basereel, vals, rhs)
    a.run(V2, e2, E_comp, basereel, vals, rhs, update_only=True)
    fv = [torch.full((c.nnz,), 3.0, dtype=dt, device="cuda") for _ in e2]
    fr = [torch.full((c.nligne, c.nlmso), 3.0, dtype=dt, device="cuda") for _ in e2]
    a.run(V2, e2, E_comp, basereel, fv, fr)
    torch.cuda.synchronize()
    for u, f in zip(vals + rhs, fv + fr):
        assert torch.equal(u, f)


def test_device_assembly_feeds_grouped_solver():
    """Energy points assembled on the device and solved in one grouped FP64
    launch sequence give the solutions of host-assembled points."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    from fdmnes_amd.batch import FP64, DeviceSolver
    c = CL.Cluster(2.4 / BOHR, 0.2 / BOHR, 3)
    lb1, lb2 = c.profile()
    rng = np.random.default_rng(6)
    ens = _energies(4, True)
    V = [rng.uniform(-1.5, 0.0, c.npoint) for _ in ens]
    a = DeviceAssembler(c)
    vals = [torch.zeros(c.nnz, dtype=torch.complex128, device="cuda") for _ in ens]
    rhs = [torch.zeros((c.nligne, c.nlmso), dtype=torch.complex128, device="cuda") for _ in ens]
    a.run([torch.from_numpy(v).cuda() for v in V], ens, True, False, vals, rhs)
    rp, col, _, _ = c.assemble(V[0], *ens[0], True, False)
    sv = DeviceSolver(FP64, True, c.nlmso)
    slots = [sv.add_system(lb1, lb2, rp, col) for _ in ens]
    xd = [torch.zeros_like(r) for r in rhs]
    sv.solve(slots, vals, rhs, xd)
    hv, hr = [], []
    for k, (E, Ei, Ec) in enumerate(ens):
        _, _, v, r = c.assemble(V[k], E, Ei, Ec, True, False)
        hv.append(torch.from_numpy(v).cuda())
        hr.append(torch.from_numpy(np.ascontiguousarray(r)).cuda())
    xh = [torch.zeros_like(r) for r in hr]
    sv.solve(slots, hv, hr, xh)
    for a_, b_ in zip(xd, xh):
        a_, b_ = a_.cpu().numpy(), b_.cpu().numpy()
        assert np.abs(a_ - b_).max() <= 1e-11 * np.abs(b_).max()


def test_device_assembly_full_size_iron_cluster():
    """Sixteen energy points of the 65,283-unknown iron cluster (R = 5 A,
    h = 0.2 A, lmax 3) in one call; two of them checked against the host."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    w = cfg.WORKLOADS["fe59_R5.0_h0.2_c128"]
    c = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    t = c.assembly_tables()
    V = cfg.potential(w, c.points)
    E = [float(e) for e in w.energies_Ry[::13][:16]]
    ens = [(e, w.eimag_Ry, cfg.ecinetic_out(e)) for e in E]
    from fdmnes_amd.assemble import DeviceAssembler
    a = DeviceAssembler(c)
    Vd = torch.from_numpy(V).cuda()
    vals = [torch.zeros(c.nnz, dtype=torch.complex128, device="cuda") for _ in ens]
    a.run([Vd] * len(ens), ens, True, False, vals)
    torch.cuda.synchronize()
    fixed = _fixed_mask(c, t)
    for k in (0, len(ens) - 1):
        _, _, hv, _ = c.assemble(V, *ens[k], True, False)
        dv = vals[k].cpu().numpy()
        assert np.array_equal(dv[fixed], hv[fixed]) and np.array_equal(dv[t["qdiag"]], hv[t["qdiag"]])
        assert np.abs(dv - hv).max() <= 1e-13 * np.abs(hv).max()


def test_assembler_rejects_bad_buffers():
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    c = CL.Cluster(1.2 / BOHR, 0.25 / BOHR, 2)
    a = DeviceAssembler(c, max_points=2)
    V = torch.zeros(c.npoint, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):  # real buffers for complex arithmetic
        a.run([V], _energies(1, True), True, False, [torch.zeros(c.nnz, dtype=torch.float64, device="cuda")])
    with pytest.raises(ValueError):  # wrong length
        a.run([V], _energies(1, True), True, False, [torch.zeros(c.nnz - 1, dtype=torch.complex128, device="cuda")])
    # more points than max_points run in several calls
    vals = [torch.zeros(c.nnz, dtype=torch.complex128, device="cuda") for _ in range(3)]
    a.run([V] * 3, _energies(3, True), True, False, vals)
    torch.cuda.synchronize()
    _, _, hv, _ = c.assemble(np.zeros(c.npoint), *_energies(3, True)[2], True, False)
    assert np.abs(vals[2].cpu().numpy() - hv).max() <= 1e-13 * np.abs(hv).max()
