"""Clusters with muffin-tin atoms on the GPU: the product's rows
(fdmgpu_cluster_assemble_atoms: grid rows coupling to an atom's expansion,
mat.f90:1233-1309, and the atom rows, :1515-1632) solved by the FP64 drop-in,
the mixed-precision drop-in and the grouped device solver, against the
reference's solution of the same cluster (tests/golden/atoms_R1.9_h0.25.npz:
three atoms, two at the same distance from the absorber, so two atom blocks
follow each other among the rows) at every third row; and the Fortran
drop-in behind the reference's own 93-argument mat_solve with the reference's
calcMatRow assembling the atom rows.  Bound 1e-10 relative to the largest
component: in the real basis the system's conditioning alone carries the
last-bit differences of the entries to 3e-11 (the oracle's elimination of the
product's rows, tests/test_cluster_atoms.py)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_atoms import BOHR, BRANCHES, G, phiato

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "fdmnes_amd", "libfdmnes_shim.so")


def _system(branch):
    from fdmnes_amd import cluster as CL
    g = dict(np.load(G))
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    c = CL.Cluster(R, h, lm, atoms=at)
    o = O.Cluster(R, h, lm, atoms=at)
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    phi = phiato(o, g["posi"] / BOHR, g["lmax"], E, Ei, E_comp)
    rp, col, vals, rhs = c.assemble_atoms(g["V"][: c.npoint], E, Ei, E, E_comp, basereel, phi)
    lb1, lb2 = c.profile()
    return g, c, o, phi, (lb1, lb2, rp, col, vals, rhs), cal


def _rel(x, g, branch):
    ref = g[f"x{branch}"]
    return np.abs(x[g["rows"]] - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("branch", range(3))
def test_dropin_with_atoms_matches_reference(branch):
    from fdmnes_amd import solver as SV
    g, c, _, _, (lb1, lb2, rp, col, vals, rhs), cal = _system(branch)
    print(f"atoms branch {BRANCHES[branch]} nligne {c.nligne}", flush=True)
    x, _, _ = SV.mat_solve(lb1, lb2, rp, col, vals, rhs, cal)
    r64 = _rel(x, g, branch)
    xm, sweeps, _ = SV.mat_solve_mixed(lb1, lb2, rp, col, vals, rhs, cal)
    r32 = _rel(xm, g, branch)
    print(f"fp64 drop-in vs reference {r64:.2e}; mixed (sweeps {sweeps}) {r32:.2e}", flush=True)
    assert r64 <= 1e-10 and r32 <= 1e-10


def test_grouped_solver_with_atoms():
    import torch
    from fdmnes_amd import batch as B
    systems = [_system(b) for b in range(3)]
    for prec in (B.FP64, B.FP32):
        sv = B.DeviceSolver(prec, True, systems[0][4][5].shape[1])
        slots = [sv.add_system(s[4][0], s[4][1], s[4][2], s[4][3]) for s in systems]
        vals = [torch.from_numpy(s[4][4].astype(np.complex128)).cuda() for s in systems]
        src = [torch.from_numpy(np.ascontiguousarray(s[4][5], np.complex128)).cuda() for s in systems]
        xs = [torch.zeros_like(t) for t in src]
        _, sweeps, _ = sv.solve(slots, vals, src, xs)
        rel = [_rel(x.cpu().numpy(), s[0], b) for b, (s, x) in enumerate(zip(systems, xs))]
        print("precision", prec, "sweeps", sweeps, "rel vs reference", rel, flush=True)
        assert max(rel) <= 1e-10
        del sv, vals, src, xs
        torch.cuda.empty_cache()


@pytest.mark.parametrize("branch", range(3))
def test_shim_with_atoms_matches_reference_mat_solve(branch):
    from oracle import ref as R
    if not (R.available() and os.path.exists(SHIM)):
        pytest.skip("needs oracle/_ref and the shim (built where the reference sources are)")
    import torch  # noqa: F401  (one HIP runtime for the process)
    from fdmnes_amd import _lib
    _lib.lib()
    R.lib()
    shim = ctypes.CDLL(SHIM, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    g, c, o, phi, _, cal = _system(branch)
    E, Ei, E_comp, _, basereel = BRANCHES[branch]
    s = R.atom_system(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]), g["posi"] / BOHR,
                      g["rmt"] / BOHR, g["lmax"])
    bes, neu = R.phiso_sys(s, E_comp, E, Ei, E)
    print(f"shim with atoms: branch {BRANCHES[branch]} nligne {s['nligne']}", flush=True)
    x_ref = R.mat_solve_sys(s, g["V"], E, Ei, bes, neu, E_comp, cal, basereel, phi)
    x_gpu = R.mat_solve_sys(s, g["V"], E, Ei, bes, neu, E_comp, cal, basereel, phi, fn=shim.mat_solve_)
    assert np.array_equal(x_ref[g["rows"]], g[f"x{branch}"])
    rel = np.abs(x_gpu - x_ref).max() / np.abs(x_ref).max()
    print("shim vs reference mat_solve", rel, flush=True)
    assert rel <= 1e-10


def _energies(E_comp):
    # (Enervide, Eimag, Ecinetic_out) in Ry
    return [(0.3 + 0.35 * k, 0.05 if E_comp else 0.0, 0.3 + 0.35 * k) for k in range(3)]


@pytest.mark.parametrize("branch", range(3))
def test_device_assembly_with_atoms_matches_host(branch):
    """fdmgpu_assembler_run_atoms at three energies, each with its own V and
    phiato, against the host assembly (fdmgpu_cluster_assemble_atoms): bit
    for bit where no outer-sphere function enters, within 1e-13 elsewhere;
    then update_only on the same buffers equals a fresh call, and the
    device-assembled systems solve to the host-assembled systems' solutions."""
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd.assemble import DeviceAssembler
    g, c, o, _, _, cal = _system(branch)
    _, _, E_comp, _, basereel = BRANCHES[branch]
    ens = _energies(E_comp)
    rng = np.random.default_rng(11 + branch)
    Vs = [g["V"][: c.npoint] + rng.uniform(-0.05, 0.05, c.npoint) for _ in ens]
    phis = [phiato(o, g["posi"] / BOHR, g["lmax"], E, Ei, E_comp) for E, Ei, _ in ens]
    print(f"device assembly with atoms: branch {BRANCHES[branch]} nnz {c.nnz}", flush=True)
    dt = torch.complex128 if cal else torch.float64
    a = DeviceAssembler(c, max_points=2)  # three points in two calls
    Vd = [torch.from_numpy(v).cuda() for v in Vs]
    pd = [torch.from_numpy(np.concatenate([p.reshape(-1) for p in ph])).cuda() for ph in phis]
    vals = [torch.full((c.nnz,), 7.0, dtype=dt, device="cuda") for _ in ens]
    rhs = [torch.full((c.nligne, c.nlmso), 7.0, dtype=dt, device="cuda") for _ in ens]
    a.run(Vd, ens, E_comp, basereel, vals, rhs, phi=pd)
    torch.cuda.synchronize()
    t = c.assembly_tables()
    fixed = np.ones(c.nnz, bool)
    for r in range(t["nrow_out"]):
        fixed[t["oq"][r]:t["oq"][r] + c.nlmso] = False
    nso = c.nligne - c.nlmso
    for lm in range(c.nlmso):
        q = t["row_ptr"][nso + lm] + c.nsortf
        fixed[q:q + c.nlmso] = False
    host = []
    for k, (E, Ei, Ec) in enumerate(ens):
        rp, col, hv, hr = c.assemble_atoms(Vs[k], E, Ei, Ec, E_comp, basereel, phis[k])
        host.append((rp, col, hv, hr))
        dv, dr = vals[k].cpu().numpy(), rhs[k].cpu().numpy()
        assert np.array_equal(dv[fixed], hv[fixed])
        assert np.abs(dv - hv).max() <= 1e-13 * np.abs(hv).max()
        assert np.abs(dr - hr).max() <= 1e-13 * np.abs(hr).max()
    # update_only: new potentials and phiato over the first call's output
    Vd2 = [torch.from_numpy(v[::-1].copy()).cuda() for v in Vs]
    a.run(Vd2, ens[::-1], E_comp, basereel, vals, rhs, update_only=True, phi=pd[::-1])
    fv = [torch.full((c.nnz,), 3.0, dtype=dt, device="cuda") for _ in ens]
    fr = [torch.full((c.nligne, c.nlmso), 3.0, dtype=dt, device="cuda") for _ in ens]
    a.run(Vd2, ens[::-1], E_comp, basereel, fv, fr, phi=pd[::-1])
    torch.cuda.synchronize()
    for u, f in zip(vals + rhs, fv + fr):
        assert torch.equal(u, f)
    # device-assembled and host-assembled points through the grouped FP64 solver
    a.run(Vd, ens, E_comp, basereel, vals, rhs, phi=pd)
    lb1, lb2 = c.profile()
    sv = B.DeviceSolver(B.FP64, True, c.nlmso)
    slots = [sv.add_system(lb1, lb2, host[0][0], host[0][1]) for _ in ens]
    cv = [v.to(torch.complex128) for v in vals]
    cr = [r.to(torch.complex128) for r in rhs]
    xd = [torch.zeros_like(r) for r in cr]
    sv.solve(slots, cv, cr, xd)
    hv = [torch.from_numpy(h[2].astype(np.complex128)).cuda() for h in host]
    hr = [torch.from_numpy(np.ascontiguousarray(h[3], np.complex128)).cuda() for h in host]
    xh = [torch.zeros_like(r) for r in hr]
    sv.solve(slots, hv, hr, xh)
    for x1, x2 in zip(xd, xh):
        x1, x2 = x1.cpu().numpy(), x2.cpu().numpy()
        rel = np.abs(x1 - x2).max() / np.abs(x2).max()
        print("device- vs host-assembled solution", rel, flush=True)
        assert rel <= 1e-10


def test_atom_assembler_needs_phiato():
    import torch
    from fdmnes_amd.assemble import DeviceAssembler
    g, c, o, _, _, _ = _system(0)
    a = DeviceAssembler(c)
    V = torch.from_numpy(g["V"][: c.npoint].copy()).cuda()
    vals = [torch.zeros(c.nnz, dtype=torch.complex128, device="cuda")]
    with pytest.raises(ValueError):
        a.run([V], _energies(True)[:1], True, False, vals)
    with pytest.raises(ValueError):  # complex phiato in real arithmetic
        a.run([V], _energies(False)[:1], False, True, [torch.zeros(c.nnz, dtype=torch.float64, device="cuda")],
              phi=[torch.zeros(a.nphi, dtype=torch.complex128, device="cuda")])
