"""The atoms' radial solutions and phiato on the device (fdmgpu_sphere_*):
for the three atoms of the atom fixture at complex and real energies, with
and without the scalar-relativistic terms, the device's phiato at every
border point and the atomic Tau match the oracle's Sphere restatement
(pinned to the reference's own Sphere, tests/test_oracle_sphere.py) within
1e-12; real output equals the complex output's real part and is refused for
complex radial solutions; and the device phiato, fed to the device assembler
of the cluster with atoms and solved by the grouped solver, gives the
solution of the oracle's whole chain (its own Sphere, rows and elimination)
within 1e-10."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_sphere import G, SPHERE, _cluster, border

pytestmark = pytest.mark.gpu
ENERGIES = [(0.7, 0.05, 1.0), (0.9, 0.0, 1.2), (0.45, 0.02, 0.6), (1.3, 0.0, 1.5)]  # Ecinetic, Eimag, Enervide


def _atoms(g, o, f):
    out = []
    for ia in range(o.natome):
        bx, pos, ylm = border(o, g, ia)
        out.append(dict(r=f[f"r{ia}"], V=f[f"V{ia}"], Z=int(f[f"Z{ia}"]), rmtg=g["rmt"][ia] / BOHR,
                        lmax=int(g["lmax"][ia]), bxyz=bx, pos=pos, ylmato=ylm))
    return out


BOHR = 0.52917721067


@pytest.mark.parametrize("rel", [False, True])
def test_device_sphere_matches_oracle(rel):
    import torch
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    print(f"device Sphere: {len(atoms)} atoms, relativistic {rel}, {len(ENERGIES)} energies", flush=True)
    ds = DeviceSphere(atoms, relativiste=rel, max_points=3)  # four points in two calls
    phi, tau = ds.run(ENERGIES, complex_out=True, with_tau=True)
    torch.cuda.synchronize()
    for k, (Ec, Ei, Ev) in enumerate(ENERGIES):
        dp, dt = phi[k].cpu().numpy(), tau[k].cpu().numpy()
        op, ot = [], []
        for a in atoms:
            p, t, _ = O.sphere(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], rel, Ec, Ei, Ev, a["bxyz"], a["pos"],
                               a["ylmato"])
            op.append(p.reshape(-1))
            ot.append(t)
        op, ot = np.concatenate(op), np.concatenate(ot)
        rp = np.abs(dp - op).max() / np.abs(op).max()
        rt = np.abs(dt - ot).max() / np.abs(ot).max()
        print(f"energy {k}: phiato rel {rp:.2e}, Tau rel {rt:.2e}", flush=True)
        assert rp <= 1e-12 and rt <= 1e-12
    # real output for the real energies
    real_e = [e for e in ENERGIES if e[1] == 0]
    pr = ds.run(real_e, complex_out=False)
    pc = ds.run(real_e, complex_out=True)
    for a, b in zip(pr, pc):
        assert torch.equal(a, b.real.contiguous())
        assert torch.all(b.imag == 0)
    from fdmnes_amd import _lib
    with pytest.raises(_lib.FdmGpuError):
        ds.run(ENERGIES[:1], complex_out=False)


def test_device_sphere_feeds_the_assembler_and_solver():
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    c = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), atoms=at)
    ens = [(1.0, 0.05, 0.7), (0.8, 0.03, 0.5)]  # (Enervide, Eimag, Ecinetic_out) for the assembler
    V = g["V"][: c.npoint]
    ds = DeviceSphere(atoms)
    phi = ds.run([(Ec, Ei, Ev) for Ev, Ei, Ec in ens], complex_out=True)
    asm = DeviceAssembler(c, max_points=2)
    vals = [torch.zeros(c.nnz, dtype=torch.complex128, device="cuda") for _ in ens]
    rhs = [torch.zeros((c.nligne, c.nlmso), dtype=torch.complex128, device="cuda") for _ in ens]
    asm.run([torch.from_numpy(V).cuda() for _ in ens], ens, True, False, vals, rhs, phi=phi)
    lb1, lb2 = c.profile()
    rp, col, _, _ = c.assemble_atoms(V, *ens[0], True, False, [np.zeros((len(a["bxyz"]), (a["lmax"] + 1) ** 2))
                                                               for a in atoms])
    sv = B.DeviceSolver(B.FP64, True, c.nlmso)
    slots = [sv.add_system(lb1, lb2, rp, col) for _ in ens]
    xs = [torch.zeros_like(r) for r in rhs]
    sv.solve(slots, vals, rhs, xs)
    for (Ev, Ei, Ec), x in zip(ens, xs):
        ophi = []
        for a in atoms:
            p, _, _ = O.sphere(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], False, Ec, Ei, Ev, a["bxyz"], a["pos"],
                               a["ylmato"])
            ophi.append(p)
        o.set_phiato(ophi)
        st, xo, *_ = o.solve(g["V"], Ev, Ei, Ec, True, True, False)
        assert st == 0
        rel = np.abs(x.cpu().numpy() - xo).max() / np.abs(xo).max()
        print(f"energy {Ev}: device Sphere + assembler + solver vs the oracle chain {rel:.2e}", flush=True)
        assert rel <= 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_energy_loop_with_atoms_and_device_sphere(precision):
    """EnergyLoop over a cluster with atoms, phiato from the device's radial
    solver per energy: solutions and tau against the oracle's chain (its
    Sphere, rows, elimination and tau reduction)."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import EnergyLoop
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    c = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), atoms=at)
    ens = [(1.0, 0.05, 0.7), (0.8, 0.03, 0.5), (1.3, 0.04, 0.9)]  # (Enervide, Eimag, Ecinetic_out)
    V = g["V"][: c.npoint]
    loop = EnergyLoop(c, precision=precision, group=2, sphere=DeviceSphere(atoms))
    out = loop.run(V, ens, keep_solutions=True)
    blocks = [(c.atom(a)["ianew"], c.atom(a)["nlmsa"]) for a in range(c.natome)]
    for k, (Ev, Ei, Ec) in enumerate(ens):
        o.set_phiato([O.sphere(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], False, Ec, Ei, Ev, a["bxyz"],
                               a["pos"], a["ylmato"])[0] for a in atoms])
        st, xo, *_ = o.solve(g["V"], Ev, Ei, Ec, True, True, False)
        assert st == 0
        rx = np.abs(out["x"][k] - xo).max() / np.abs(xo).max()
        to = O.tau(xo, o.nligne - o.nlmso, False, False, blocks)[0]
        tg = out["tau"][k]
        rt = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(tg, to))
        print(f"precision {precision} energy {k}: solution rel {rx:.2e}, tau rel {rt:.2e}", flush=True)
        assert rx <= 1e-10 and rt <= 1e-9


@pytest.mark.parametrize("rel", [False, True])
def test_device_irregular_solution_matches_oracle(rel):
    """fdmgpu_sphere_run_ri: the regular and the irregular solutions'
    phiato against the oracle's Sphere with n_RI = 2 (pinned to the
    reference's, tests/test_oracle_sphere.py) within 1e-12."""
    import torch
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    print(f"device irregular Sphere: relativistic {rel}, {len(ENERGIES)} energies", flush=True)
    ds = DeviceSphere(atoms, relativiste=rel, max_points=3)
    out = ds.run_ri(ENERGIES, complex_out=True)
    torch.cuda.synchronize()
    for k, (Ec, Ei, Ev) in enumerate(ENERGIES):
        o1, o2 = [], []
        for a in atoms:
            p, p2, _, _ = O.sphere_ri(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], rel, Ec, Ei, Ev, a["bxyz"],
                                      a["pos"], a["ylmato"])
            o1.append(p.reshape(-1))
            o2.append(p2.reshape(-1))
        o1, o2 = np.concatenate(o1), np.concatenate(o2)
        d1, d2 = out[k][0].cpu().numpy(), out[k][1].cpu().numpy()
        r1 = np.abs(d1 - o1).max() / np.abs(o1).max()
        r2 = np.abs(d2 - o2).max() / np.abs(o2).max()
        print(f"energy {k}: regular rel {r1:.2e}, irregular rel {r2:.2e}", flush=True)
        assert r1 <= 1e-12 and r2 <= 1e-12


def test_device_irregular_phiato_assembles_the_irregular_system():
    """The device's regular and irregular phiato, given to the host
    assembler of the irregular cluster system (calcMatRow's irregular
    branches), write the rows the oracle's Sphere values write (1e-12)."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    ci = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), atoms=at, irregular=True)
    Ev, Ei, Ec = 1.0, 0.05, 0.7
    out = DeviceSphere(atoms).run_ri([(Ec, Ei, Ev)])
    torch.cuda.synchronize()
    d1, d2 = out[0][0].cpu().numpy(), out[0][1].cpu().numpy()
    blocks = lambda flat: [flat[s:s + len(a["bxyz"]) * (a["lmax"] + 1) ** 2].reshape(len(a["bxyz"]), -1)
                           for s, a in zip(np.cumsum([0] + [len(a["bxyz"]) * (a["lmax"] + 1) ** 2
                                                            for a in atoms])[:-1], atoms)]
    op = [O.sphere_ri(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], False, Ec, Ei, Ev, a["bxyz"], a["pos"],
                      a["ylmato"]) for a in atoms]
    V = g["V"][: ci.npoint]
    _, _, vd, rd = ci.assemble_atoms(V, Ev, Ei, Ec, True, False, blocks(d1), blocks(d2))
    _, _, vo, ro = ci.assemble_atoms(V, Ev, Ei, Ec, True, False, [p[0] for p in op], [p[1] for p in op])
    assert np.abs(vd - vo).max() <= 1e-12 * np.abs(vo).max()
    assert np.abs(rd - ro).max() <= 1e-12 * max(np.abs(ro).max(), 1e-300)


def test_device_assembly_of_the_irregular_system():
    """fdmgpu_assembler_run_ri: the device assembler of the irregular cluster
    solution (two blocks per atom, each atom row block on its half of the
    border, no outer unknowns, incoming wave -i j + n), fed the device's
    regular and irregular phiato, writes the host assembler's rows: fixed
    entries bit for bit, the rest within 1e-13; fresh and updated in place."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    ci = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), atoms=at, irregular=True)
    ens = [(1.0, 0.05, 0.7), (0.8, 0.03, 0.5)]  # (Enervide, Eimag, Ecinetic_out)
    V = g["V"][: ci.npoint]
    pp = DeviceSphere(atoms).run_ri([(Ec, Ei, Ev) for Ev, Ei, Ec in ens])
    phi, phi2 = [p[0] for p in pp], [p[1] for p in pp]
    asm = DeviceAssembler(ci, max_points=2)
    vals = [torch.full((ci.nnz,), 5.0, dtype=torch.complex128, device="cuda") for _ in ens]
    rhs = [torch.full((ci.nligne, ci.nlmso), 5.0, dtype=torch.complex128, device="cuda") for _ in ens]
    Vd = [torch.from_numpy(np.ascontiguousarray(V)).cuda() for _ in ens]
    asm.run(Vd, ens, True, False, vals, rhs, phi=phi, phi2=phi2)
    asm.run(Vd[::-1], ens[::-1], True, False, vals, rhs, update_only=True, phi=phi[::-1], phi2=phi2[::-1])
    torch.cuda.synchronize()
    offs = np.cumsum([0] + [len(a["bxyz"]) * (a["lmax"] + 1) ** 2 for a in atoms])
    blocks = lambda t: [t.cpu().numpy()[offs[k]:offs[k + 1]].reshape(len(a["bxyz"]), -1) for k, a in enumerate(atoms)]
    for k, e in enumerate(ens[::-1]):
        kk = len(ens) - 1 - k
        _, _, hv, hr = ci.assemble_atoms(V, *e, True, False, blocks(phi[kk]), blocks(phi2[kk]))
        dv, dr = vals[k].cpu().numpy(), rhs[k].cpu().numpy()
        err_v = np.abs(dv - hv).max() / np.abs(hv).max()
        err_r = np.abs(dr - hr).max() / np.abs(hr).max()
        print(f"irregular system energy {k}: rows rel {err_v:.2e}, rhs rel {err_r:.2e}", flush=True)
        assert err_v <= 1e-13 and err_r <= 1e-13


def test_fdm_comp_loop_gives_the_host_chains_tau_ri():
    """FdmCompLoop (device radial solver with the irregular solution, device
    assembly of the regular and irregular systems, one grouped solve, device
    Tau_RI) against the same chain with the oracle's Sphere, the host
    assemblers and the FP64 grouped solve, three complex energies (1e-9)."""
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import tau as T
    from fdmnes_amd.energy_loop import FdmCompLoop
    from fdmnes_amd.sphere import DeviceSphere
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    c = CL.Cluster(R, h, lm, atoms=at)
    ci = CL.Cluster(R, h, lm, atoms=at, irregular=True)
    ens = [(1.0, 0.05, 0.7), (0.8, 0.03, 0.5), (1.3, 0.04, 0.9)]
    V = g["V"][: c.npoint]
    out = FdmCompLoop(c, ci, DeviceSphere(atoms), group=2).run(V, ens)
    nl = c.nlmso
    for k, (Ev, Ei, Ec) in enumerate(ens):
        op = [O.sphere_ri(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], False, Ec, Ei, Ev, a["bxyz"], a["pos"],
                          a["ylmato"]) for a in atoms]
        phi, phi2 = [p[0] for p in op], [p[1] for p in op]
        reg = c.assemble_atoms(V, Ev, Ei, Ec, True, False, phi)
        irr = ci.assemble_atoms(V, Ev, Ei, Ec, True, False, phi, phi2)
        sv = B.DeviceSolver(B.FP64, True, nl)
        slots = [sv.add_system(*c.profile(), reg[0], reg[1]), sv.add_system(*ci.profile(), irr[0], irr[1])]
        vals = [torch.from_numpy(np.ascontiguousarray(s[2], np.complex128)).cuda() for s in (reg, irr)]
        src = [torch.from_numpy(np.ascontiguousarray(s[3], np.complex128)).cuda() for s in (reg, irr)]
        xs = [torch.zeros_like(s) for s in src]
        sv.solve(slots, vals, src, xs)
        ar = [(c.atom(a)["ianew"], c.atom(a)["nlmsa"]) for a in range(c.natome)]
        ai = [(ci.atom(a)["ianew"], ci.atom(a)["nlmsa"]) for a in range(ci.natome)]
        want = T.tau_ri([xs[0]], [xs[1]], ar, ai)[0].cpu().numpy()
        got = np.concatenate([b.reshape(-1) for b in out["tau_ri"][k]])
        rel = np.abs(got - want).max() / np.abs(want).max()
        print(f"energy {k}: FdmCompLoop Tau_RI vs the host chain {rel:.2e}", flush=True)
        assert rel <= 1e-9
        del sv
