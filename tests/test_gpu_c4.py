"""Complex harmonics (Ylm_comp) and the four-fold groups 4, -4, 4/m on the
GPU: every representation of each group at two energies in one grouped
solve, FP64 and FP32 factors with FP64 refinement, reproduces the oracle's
whole-cluster solution in complex harmonics for a potential with exactly the
group's symmetry (bound 1e-10, the solver's parity bar); and the P1 system
with complex harmonics through the FP64 and mixed drop-ins reproduces the
reference's own solution (tests/golden/ylmc_R1.2_h0.25.npz, mat_solve with
Ylm_comp from oracle/_ref)."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_cluster_c4 import BOHR, IRREPS, YLM_BRANCHES, YLMC, potential
from .test_gpu_mmm import _solve_reduced

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("group", [9, 10, 11])
def test_grouped_complex_representations_match_oracle(group, precision):
    from fdmnes_amd import cluster as CL
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    reps = [CL.Cluster(R, h, lm, irrep=v, group=group) for v in IRREPS[group]]
    ens = [(0.35, 0.0367, 0.55), (0.8, 0.0367, 1.0)]
    print(f"group {group} precision {precision}: {len(reps)} representations x {len(ens)} energies, rows",
          [c.nligne for c in reps], flush=True)
    out = _solve_reduced(reps, lambda p: potential(p, group), ens, True, False, precision)
    o = O.Cluster(R, h, lm)
    o.set_ylm_comp(True)
    xyz = o.xyz[:o.npoint, :3]
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    nsof = o.nligne - o.nlmso
    full = {}
    for e in ens:
        st, xf, *_ = o.solve(potential(xyz, group), *e, True, True, False)
        assert st == 0
        full[e] = xf
    for c, e, x in out:
        xf = full[e]
        hm = c.harmonics
        rows = [row[tuple(k)] for k in np.round(c.points / h).astype(int)]
        d = max(np.abs(x[:c.npoint, :c.nlmso] / c.rvol[:, None] - xf[rows][:, hm]).max(),
                np.abs(x[c.npoint:, :c.nlmso] - xf[nsof + hm][:, hm]).max()) / np.abs(xf).max()
        print("irrep", c.irrep, "energy", e[0], "rel", d, flush=True)
        assert d <= 1e-10


@pytest.mark.parametrize("branch", range(3))
def test_ylm_comp_dropin_matches_reference(branch):
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import solver as SV
    g = np.load(YLMC)
    c = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), ylm_comp=True)
    E, Ei, E_comp, basereel = YLM_BRANCHES[branch]
    rp, col, vals, rhs = c.assemble(g["V"], E, Ei, E, E_comp, basereel)
    lb1, lb2 = c.profile()
    ref = g[f"x{branch}"]
    x, _, _ = SV.mat_solve(lb1, lb2, rp, col, vals, rhs, True)
    xm, sweeps, _ = SV.mat_solve_mixed(lb1, lb2, rp, col, vals, rhs, True)
    r64 = np.abs(x - ref).max() / np.abs(ref).max()
    r32 = np.abs(xm - ref).max() / np.abs(ref).max()
    print(f"Ylm_comp branch {YLM_BRANCHES[branch]}: fp64 drop-in vs reference {r64:.2e}, mixed ({sweeps}) {r32:.2e}",
          flush=True)
    assert r64 <= 1e-10 and r32 <= 1e-10


@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_device_assembly_with_complex_harmonics_and_characters(E_comp, basereel):
    """fdmgpu_assembler_* with the complex tables (fdmgpu_assembler_set_
    complex) writes the host assembler's rows for every representation of
    4, -4, 4/m and for P1 in complex harmonics: energy-independent entries
    bit for bit, diagonals to 1e-15, outer-sphere entries and RHS to 1e-13
    (the device's Bessel/Neumann values), fresh and updated in place."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.assemble import DeviceAssembler
    R, h, lm = 1.6 / BOHR, 0.25 / BOHR, 2
    ens = [(0.45, 0.0367 if E_comp else 0.0, 0.7), (0.9, 0.0367 if E_comp else 0.0, 1.2)]
    cls = [CL.Cluster(R, h, lm, irrep=v, group=g) for g in (9, 10, 11) for v in IRREPS[g]
           if (g, v) != (11, 6)]  # 4/m's representation 6 has no harmonic with l <= 2
    cls.append(CL.Cluster(R, h, lm, ylm_comp=True))
    for c in cls:
        print(f"device assembly group {c.group} irrep {c.irrep} rows {c.nligne}", flush=True)
        t = c.assembly_tables()
        V = potential(c.points, c.group if c.irrep >= 0 else 9)
        a = DeviceAssembler(c)
        vals = [torch.full((c.nnz,), 5.0, dtype=torch.complex128, device="cuda") for _ in ens]
        rhs = [torch.full((c.nligne, c.nlmso), 5.0, dtype=torch.complex128, device="cuda") for _ in ens]
        Vd = [torch.from_numpy(np.ascontiguousarray(V)).cuda()] * len(ens)
        a.run(Vd, ens, E_comp, basereel, vals, rhs)
        a.run(Vd, ens[::-1], E_comp, basereel, vals, rhs, update_only=True)
        torch.cuda.synchronize()
        fixed = np.ones(c.nnz, bool)
        fixed[t["qdiag"]] = False
        for r in range(t["nrow_out"]):
            fixed[t["oq"][r]: t["oq"][r] + c.nlmso] = False
        for k in range(c.nlmso):
            q = t["row_ptr"][c.npoint + k] + c.nsortf
            fixed[q: q + c.nlmso] = False
        for k, e in enumerate(ens[::-1]):
            _, _, hv, hr = c.assemble(V, *e, E_comp, basereel)
            dv, dr = vals[k].cpu().numpy(), rhs[k].cpu().numpy()
            assert np.array_equal(dv[fixed], hv[fixed])
            assert np.abs(dv[t["qdiag"]] - hv[t["qdiag"]]).max() <= 1e-15 * np.abs(hv[t["qdiag"]]).max()
            assert np.abs(dv - hv).max() <= 1e-13 * np.abs(hv).max()
            assert np.abs(dr - hr).max() <= 1e-13 * np.abs(hr).max()


@pytest.mark.parametrize("precision", [0, 1])
def test_representation_loop_complex_matches_oracle(precision):
    """RepresentationLoop over the eight representations of 4/m (device
    assembly with complex characters and harmonics, grouped solves of 32
    systems) reproduces the oracle's whole-cluster solution in complex
    harmonics at three energies (bound 1e-10)."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import RepresentationLoop
    group = 11
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    reps = [CL.Cluster(R, h, lm, irrep=v, group=group) for v in IRREPS[group]]
    ens = [(0.35 + 0.3 * k, 0.0367, 0.55 + 0.3 * k) for k in range(3)]
    loop = RepresentationLoop(reps, precision=precision)
    out = loop.run([potential(c.points, group) for c in reps], ens, keep_solutions=True)
    o = O.Cluster(R, h, lm)
    o.set_ylm_comp(True)
    xyz = o.xyz[:o.npoint, :3]
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    nsof = o.nligne - o.nlmso
    for k, e in enumerate(ens):
        st, xf, *_ = o.solve(potential(xyz, group), *e, True, True, False)
        assert st == 0
        for r, c in enumerate(reps):
            x = out["x"][k][r]
            hm = c.harmonics
            rows = [row[tuple(q)] for q in np.round(c.points / h).astype(int)]
            d = max(np.abs(x[:c.npoint] / c.rvol[:, None] - xf[rows][:, hm]).max(),
                    np.abs(x[c.npoint:] - xf[nsof + hm][:, hm]).max()) / np.abs(xf).max()
            print("precision", precision, "energy", k, "irrep", c.irrep, "rel", d, flush=True)
            assert d <= 1e-10
