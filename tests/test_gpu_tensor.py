"""The absorbing atom's cross-section tensors on the device: the radial
integrals of the golden rule computed beside phiato (fdmgpu_sphere_run_
integrals) against the oracle's Radial_integral restatement within 1e-11,
and the tensors (fdmgpu_tensor_run: E1E1, E1E2, E2E2 of K, L3 core-resolved
and L23 edges, no symmetry and mmm, rotated frames, hydrogen) from those
integrals against the reference's own Tenseur_car (tests/golden/
tensor_R1.9.npz) within 1e-10 of each tensor's largest element; and the
whole chain in the energy loop (device Sphere with integrals -> assembly ->
solve -> tau -> symmetrisation -> tensors) against the oracle's chain
within 1e-9."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_tensor import TENSOR, _case

pytestmark = pytest.mark.gpu


def _device_rof(f, k):
    import torch
    from fdmnes_amd.sphere import DeviceSphere
    c, lmax, js, ls, nb, cr, rel, Z, tabs = _case(f, k)
    Ec, Ei = f[c + "E"]
    nlm = (lmax + 1) ** 2
    ds = DeviceSphere([dict(r=f["r"], V=f["V"], Z=Z, rmtg=float(f["rmt"]), lmax=lmax, bxyz=np.zeros((0, 3)),
                            pos=np.zeros(3), ylmato=np.zeros((0, nlm)))], relativiste=rel)
    ds.set_integrals(0, f[c + "psii"], f[c + "Eseuil"], float(f["rmtsd"]))
    _, rof = ds.run([(Ec, Ei, Ec + 0.3)], complex_out=True, energ=[Ec])
    torch.cuda.synchronize()
    return ds, rof[0]


@pytest.mark.parametrize("k", range(6))
def test_device_radial_integrals_match_oracle(k):
    f = np.load(TENSOR)
    c, lmax, js, ls, nb, cr, rel, Z, tabs = _case(f, k)
    print(f"radial integrals case {k}: lmax {lmax} edges {nb} relativistic {rel} Z {Z}", flush=True)
    _, rof = _device_rof(f, k)
    Ec, Ei = f[c + "E"]
    want, _ = O.radial_integral(f["r"], f["V"], Z, float(f["rmt"]), float(f["rmtsd"]), lmax, rel, Ec, Ei, Ec + 0.3,
                                Ec, f[c + "Eseuil"], f[c + "psii"])
    got = rof.cpu().numpy()
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()  # the wave sums in another order


@pytest.mark.parametrize("k", range(6))
def test_device_tensors_match_reference(k):
    import torch
    from fdmnes_amd.tensor import DeviceTensor
    f = np.load(TENSOR)
    c, lmax, js, ls, nb, cr, rel, Z, tabs = _case(f, k)
    print(f"tensors case {k}: edge ({js}, {ls}, {nb}) core resolved {cr} lmax {lmax}", flush=True)
    ds, rof = _device_rof(f, k)
    tl = f[c + "taull"]
    nlm = tl.shape[0]
    dt = DeviceTensor(js, ls, nb, lmax, tabs, ds.rof_layout, 0, nlm, core_resolved=cr, numat=Z,
                      rot_int=f[c + "rot_int"], rot_atom_abs=f[c + "rot_atom_abs"])
    tau = torch.from_numpy(np.ascontiguousarray(tl).reshape(-1)).cuda()
    dd, dq, qq = dt.run([rof, rof], [tau, tau])  # two points, the same inputs
    torch.cuda.synchronize()
    for name, got in (("dd", dd), ("dq", dq), ("qq", qq)):
        want = f[c + name]
        scale = np.abs(want).max()
        for p in range(2):
            g = got[p].cpu().numpy()
            assert g.shape == want.shape
            err = np.abs(g - want).max()
            print(f"  sec{name}: max {scale:.3e} error {err:.2e}", flush=True)
            assert err <= 1e-10 * max(scale, 1e-300)


@pytest.mark.parametrize("precision", [0, 1])
def test_energy_loop_tensors_match_oracle_chain(precision):
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import EnergyLoop
    from fdmnes_amd.sphere import DeviceSphere
    from fdmnes_amd.tensor import DeviceTensor, no_symmetry_tables

    from .test_gpu_sphere import BOHR, _atoms
    from .test_oracle_sphere import SPHERE, _cluster
    g, o = _cluster()
    f = np.load(SPHERE)
    atoms = _atoms(g, o, f)
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    c = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), atoms=at)
    a0 = atoms[0]
    lmax = int(a0["lmax"])
    psii = np.stack([a0["r"] * np.exp(-a0["Z"] * a0["r"])])
    Es, rmtsd = np.array([520.0]), 1.1 * a0["rmtg"]
    ds = DeviceSphere(atoms)
    ds.set_integrals(0, psii, Es, rmtsd)
    tabs = no_symmetry_tables()
    nl0 = (lmax + 1) ** 2
    dt = DeviceTensor(1, 0, 1, lmax, tabs, ds.rof_layout, 0, nl0, numat=int(a0["Z"]))
    ens = [(1.0, 0.05, 0.7), (0.8, 0.03, 0.5), (1.3, 0.04, 0.9)]  # (Enervide, Eimag, Ecinetic_out)
    energ = [0.6, 0.4, 0.8]
    loop = EnergyLoop(c, precision=precision, group=2, sphere=ds, tensor=dt)
    out = loop.run(g["V"][: c.npoint], ens, energ=energ)
    blocks = [(c.atom(a)["ianew"], c.atom(a)["nlmsa"]) for a in range(c.natome)]
    cg, mg, _ = O.coef_init(1, 0, 1)
    for k, (Ev, Ei, Ec) in enumerate(ens):
        o.set_phiato([O.sphere(a["r"], a["V"], a["Z"], a["rmtg"], a["lmax"], False, Ec, Ei, Ev, a["bxyz"],
                               a["pos"], a["ylmato"])[0] for a in atoms])
        st, xo, *_ = o.solve(g["V"], Ev, Ei, Ec, True, True, False)
        assert st == 0
        taus = O.tau_sym(O.tau(xo, o.nligne - o.nlmso, False, False, blocks)[0])
        rof, _ = O.radial_integral(a0["r"], a0["V"], a0["Z"], a0["rmtg"], rmtsd, lmax, False, Ec, Ei, Ev, energ[k],
                                   Es, psii)
        rofl = np.repeat(rof, [2 * L + 1 for L in range(lmax + 1)], axis=2)
        want = O.tenseur(rofl, taus[0], 0, cg, mg, 2, 1, lmax, tabs, numat=int(a0["Z"]))
        for name, w in zip(("secdd", "secdq", "secqq"), want):
            err = np.abs(out[name][k] - w).max() / np.abs(w).max()
            print(f"precision {precision} energy {k} {name}: rel {err:.2e}", flush=True)
            assert err <= 1e-9
