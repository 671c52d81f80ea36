"""The rest of the post-solve chain on the GPU (fdmgpu_tau_ri,
fdmgpu_tau_symmetrize) and the irregular cluster solution through the
device solvers:

- Tau_RI (mat.f90:229-295) on the device equals the oracle's bit for bit on
  the same solutions (fresh and accumulated), the regular solution entering
  rounded to single precision as mat.f90:211 stores it;
- the symmetrisation ending mat (:596-641) equals the oracle's bit for bit in
  both branches (real and complex harmonics);
- the irregular system of the three-atom cluster (the product's rows,
  fdmgpu_cluster_assemble_atoms_ri) solved by the FP64 drop-in and by the
  grouped solver in both precisions gives the reference's irregular solution
  (tests/golden/atoms_tau_R1.9_h0.25.npz, from oracle/_ref) within 1e-10;
- the whole chain on the device (regular and irregular solves, Tau_RI) gives
  the reference mat's Tau_RI within 1e-10, and the regular Taull chain with
  the symmetrisation gives mat's Taull (real and complex basis)."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_tau_ri import BOHR, _load, _split

pytestmark = pytest.mark.gpu


def _rand_solutions(rng, nligne, nlmso):
    return rng.normal(size=(nligne, nlmso)) + 1j * rng.normal(size=(nligne, nlmso))


@pytest.mark.parametrize("accumulate", [False, True])
def test_tau_ri_equals_oracle_bit_for_bit(accumulate):
    import torch
    from fdmnes_amd import tau as T
    rng = np.random.default_rng(11)
    nlmsa = [9, 4, 16, 1]
    nlmso = 9
    reg = [(5, 9), (40, 4), (60, 16), (90, 1)]
    irr = [(0, 9), (30, 4), (50, 16), (100, 1)]
    xr = [_rand_solutions(rng, 120, nlmso) for _ in range(3)]
    xi = [_rand_solutions(rng, 140, nlmso) for _ in range(3)]
    tot = sum(2 * k * k for k in nlmsa)
    t0 = [rng.normal(size=tot) + 1j * rng.normal(size=tot) for _ in range(3)]
    print("tau_ri points 3, atoms", nlmsa, "accumulate", accumulate, flush=True)
    dev = [torch.from_numpy(x).cuda() for x in xr], [torch.from_numpy(x).cuda() for x in xi]
    tau = [torch.from_numpy(t.copy()).cuda() for t in t0] if accumulate else None
    out = T.tau_ri(dev[0], dev[1], reg, irr, tau=tau, accumulate=accumulate)
    torch.cuda.synchronize()
    for k in range(3):
        blocks, off = [], 0
        for na in nlmsa:
            blocks.append(t0[k][off:off + 2 * na * na].reshape(2, na, na))
            off += 2 * na * na
        ref = O.tau_ri(xr[k], xi[k], [a for a, _ in reg], [a for a, _ in irr], nlmsa,
                       tau0=blocks if accumulate else None)
        got = out[k].cpu().numpy()
        assert np.array_equal(got, np.concatenate([r.reshape(-1) for r in ref])), k


@pytest.mark.parametrize("ylm_comp", [False, True])
def test_tau_symmetrize_equals_oracle_bit_for_bit(ylm_comp):
    import torch
    from fdmnes_amd import tau as T
    rng = np.random.default_rng(12)
    nlmsa = [9, 16, 4, 1, 25]
    blocks = [[rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k)) for k in nlmsa] for _ in range(2)]
    tau = [torch.from_numpy(np.concatenate([b.reshape(-1) for b in bl])).cuda() for bl in blocks]
    T.symmetrize(tau, nlmsa, ylm_comp)
    torch.cuda.synchronize()
    for k in range(2):
        ref = O.tau_sym(blocks[k], ylm_comp)
        assert np.array_equal(tau[k].cpu().numpy(), np.concatenate([r.reshape(-1) for r in ref]))


def _systems():
    from fdmnes_amd import cluster as CL
    g, t = _load()
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    c = CL.Cluster(R, h, lm, atoms=at)
    ci = CL.Cluster(R, h, lm, atoms=at, irregular=True)
    o = O.Cluster(R, h, lm, atoms=at)
    phi, phi2 = _split(t["phi_c"], o), _split(t["phi2_c"], o)
    E, Ei = float(t["Ec"]), float(t["Eimag"])
    V = t["V"][: c.npoint]
    reg = c.assemble_atoms(V, E, Ei, E, True, False, phi)
    irr = ci.assemble_atoms(V, E, Ei, E, True, False, phi, phi2)
    return g, t, c, ci, reg, irr


def _rel_irr(x, t):
    ref = t["x_irr"]
    return np.abs(x[t["irr_rows"]] - ref).max() / np.abs(ref).max()


def test_irregular_system_solves_to_reference():
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import solver as SV
    g, t, c, ci, reg, irr = _systems()
    lb1, lb2 = ci.profile()
    rp, col, vals, rhs = irr
    print(f"irregular system nligne {ci.nligne}", flush=True)
    x, _, _ = SV.mat_solve(lb1, lb2, rp, col, vals, rhs, True)
    r64 = _rel_irr(x, t)
    for prec in (B.FP64, B.FP32):
        sv = B.DeviceSolver(prec, True, rhs.shape[1])
        slot = sv.add_system(lb1, lb2, rp, col)
        v = torch.from_numpy(vals.astype(np.complex128)).cuda()
        s = torch.from_numpy(np.ascontiguousarray(rhs, np.complex128)).cuda()
        xg = torch.zeros_like(s)
        _, sweeps, _ = sv.solve([slot], [v], [s], [xg])
        rel = _rel_irr(xg.cpu().numpy(), t)
        print(f"drop-in {r64:.2e}; grouped precision {prec} sweeps {sweeps} {rel:.2e}", flush=True)
        assert rel <= 1e-10
        del sv
    assert r64 <= 1e-10


def test_chain_on_device_gives_reference_tau_ri_and_taull():
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import tau as T
    g, t, c, ci, reg, irr = _systems()
    nl = reg[3].shape[1]
    sv = B.DeviceSolver(B.FP64, True, nl)
    slots = [sv.add_system(*c.profile(), reg[0], reg[1]), sv.add_system(*ci.profile(), irr[0], irr[1])]
    vals = [torch.from_numpy(s[2].astype(np.complex128)).cuda() for s in (reg, irr)]
    src = [torch.from_numpy(np.ascontiguousarray(s[3], np.complex128)).cuda() for s in (reg, irr)]
    xs = [torch.zeros_like(s) for s in src]
    sv.solve(slots, vals, src, xs)
    atoms_reg = [(c.atom(a)["ianew"], c.atom(a)["nlmsa"]) for a in range(c.natome)]
    atoms_irr = [(ci.atom(a)["ianew"], ci.atom(a)["nlmsa"]) for a in range(ci.natome)]
    out = T.tau_ri([xs[0]], [xs[1]], atoms_reg, atoms_irr)[0].cpu().numpy()
    xr, xi = xs[0].cpu().numpy(), xs[1].cpu().numpy()
    ora = O.tau_ri(xr, xi, [a for a, _ in atoms_reg], [a for a, _ in atoms_irr], [k for _, k in atoms_reg])
    assert np.array_equal(out, np.concatenate([r.reshape(-1) for r in ora]))
    ref = t["tau_ri"]
    scale = np.abs(ref).max()
    off, worst = 0, 0.0
    for a, (_, na) in enumerate(atoms_reg):
        blk = out[off:off + 2 * na * na].reshape(2, na, na)
        worst = max(worst, np.abs(blk - ref[a][:, :na, :na]).max() / scale)
        off += 2 * na * na
    print(f"device Tau_RI vs the reference's mat {worst:.2e}", flush=True)
    assert worst <= 1e-10
    del sv
    # the regular Taull chain at real energy, both bases, with the symmetrisation
    o = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]),
                  atoms=(g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"]))
    phir = _split(t["phi_real"], o)
    E = float(t["E"])
    for basereel in (True, False):
        rp, col, v, rhs = c.assemble_atoms(t["V"][: c.npoint], E, 0.0, E, False, basereel, phir)
        cal = not basereel
        sv = B.DeviceSolver(B.FP64, cal, nl)
        slot = sv.add_system(*c.profile(), rp, col)
        dt = torch.complex128 if cal else torch.float64
        vv = torch.from_numpy(np.ascontiguousarray(v)).to(dt).cuda()
        ss = torch.from_numpy(np.ascontiguousarray(rhs)).to(dt).cuda()
        xx = torch.zeros_like(ss)
        sv.solve([slot], [vv], [ss], [xx], max_sweeps=2)
        taus, _, _ = T.tau_reduce([xx], c.nligne - nl, basereel, False, atoms_reg)
        T.symmetrize(taus, [k for _, k in atoms_reg], False)
        got = taus[0].cpu().numpy()
        ref = t[f"taull_basereel{int(basereel)}"]
        off, worst = 0, 0.0
        for a, (_, na) in enumerate(atoms_reg):
            worst = max(worst, np.abs(got[off:off + na * na].reshape(na, na) - ref[a][:na, :na]).max())
            off += na * na
        print(f"device Taull (basereel {basereel}) vs the reference's mat {worst / np.abs(ref).max():.2e}", flush=True)
        assert worst <= 1e-10 * np.abs(ref).max()
        del sv
