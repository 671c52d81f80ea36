"""Scalar-relativistic rows (the reference's Relativiste = T; calcMatRow,
mat.f90:1155-1156, :1171-1185) on the GPU:

- the device assembler with relativistic tables (fdmgpu_assembler_run_rel)
  against the host assembler at three energies with their own potentials and
  gradients: bit for bit in the grid rows' interior entries and diagonals,
  within 1e-13 where the outer-sphere functions enter; update_only equals a
  fresh call; for the whole cluster, a representation of mmm and a cluster
  with atoms;
- the device-assembled relativistic system through the FP64 drop-in and the
  mixed-precision drop-in against the reference's own solution
  (tests/golden/rel_R1.2_h0.25.npz, mat_solve with Relativiste from
  oracle/_ref), bound 1e-10;
- the Fortran drop-in behind the reference's 93-argument mat_solve with
  Relativiste = T (the reference's calcMatRow assembling the rows) against
  the reference's mat_solve."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_atoms import BRANCHES, G as ATOMS_G, phiato
from .test_relativistic import ATOMS, BOHR, P1

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "fdmnes_amd", "libfdmnes_shim.so")


def _masks(c):
    """Positions the relativistic kernels write without outer-sphere
    functions (interior entries and diagonals), and the rest."""
    t, rt = c.assembly_tables(), c.rel_tables()
    exact = np.zeros(c.nnz, bool)
    exact[t["qdiag"][t["qdiag"] >= 0]] = True
    exact[rt["rq"][rt["rq"] >= 0]] = True
    return exact


def _device_vs_host(c, E_comp, basereel, host, phis=None, seed=0):
    import torch
    from fdmnes_amd.assemble import DeviceAssembler
    cal = E_comp or not basereel
    dt = torch.complex128 if cal else torch.float64
    rng = np.random.default_rng(seed)
    ens = [(0.3 + 0.35 * k, 0.05 if E_comp else 0.0, 0.3 + 0.35 * k) for k in range(3)]
    Vs = [rng.uniform(-1.0, 0.0, c.npoint) for _ in ens]
    Gs = [rng.normal(size=(c.npoint, 3)) * 30.0 for _ in ens]
    a = DeviceAssembler(c, max_points=2, relativistic=True)  # three points in two calls
    Vd = [torch.from_numpy(v).cuda() for v in Vs]
    Gd = [torch.from_numpy(np.ascontiguousarray(g)).cuda() for g in Gs]
    pd = None
    if phis is not None:
        pd = [torch.from_numpy(np.concatenate([p.reshape(-1) for p in ph])).cuda() for ph in phis]
    vals = [torch.full((c.nnz,), 7.0, dtype=dt, device="cuda") for _ in ens]
    rhs = [torch.full((c.nligne, c.nlmso), 7.0, dtype=dt, device="cuda") for _ in ens]
    a.run(Vd, ens, E_comp, basereel, vals, rhs, phi=pd, G=Gd)
    torch.cuda.synchronize()
    exact = _masks(c)
    for k, (E, Ei, Ec) in enumerate(ens):
        _, _, hv, hr = host(Vs[k], Gs[k], E, Ei, Ec, None if phis is None else phis[k])
        dv, dr = vals[k].cpu().numpy(), rhs[k].cpu().numpy()
        print(f"point {k}: interior/diagonal equal {np.array_equal(dv[exact], hv[exact])}, "
              f"max rel {np.abs(dv - hv).max() / np.abs(hv).max():.2e}", flush=True)
        assert np.array_equal(dv[exact], hv[exact])
        assert np.abs(dv - hv).max() <= 1e-13 * np.abs(hv).max()
        assert np.abs(dr - hr).max() <= 1e-13 * np.abs(hr).max()
    # update_only over the first call's output equals a fresh call
    Vd2, Gd2 = Vd[::-1], Gd[::-1]
    pd2 = pd[::-1] if pd is not None else None
    a.run(Vd2, ens[::-1], E_comp, basereel, vals, rhs, update_only=True, phi=pd2, G=Gd2)
    fv = [torch.full((c.nnz,), 3.0, dtype=dt, device="cuda") for _ in ens]
    fr = [torch.full((c.nligne, c.nlmso), 3.0, dtype=dt, device="cuda") for _ in ens]
    a.run(Vd2, ens[::-1], E_comp, basereel, fv, fr, phi=pd2, G=Gd2)
    torch.cuda.synchronize()
    for u, f in zip(vals + rhs, fv + fr):
        assert torch.equal(u, f)
    # the gradient is required exactly when the assembler is relativistic
    with pytest.raises(ValueError):
        a.run(Vd, ens, E_comp, basereel, vals, rhs, phi=pd)
    plain = DeviceAssembler(c, max_points=2)
    with pytest.raises(ValueError):
        plain.run(Vd[:1], ens[:1], E_comp, basereel, vals[:1], rhs[:1], phi=pd[:1] if pd else None, G=Gd[:1])


@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True), (False, False)])
def test_device_relativistic_assembly_matches_host(E_comp, basereel):
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(1.6 / BOHR, 0.3 / BOHR, 2)
    print(f"relativistic device assembly P1 n={c.nligne} E_comp={E_comp} basereel={basereel}", flush=True)
    _device_vs_host(c, E_comp, basereel,
                    lambda V, G, E, Ei, Ec, _: c.assemble(V, E, Ei, Ec, E_comp, basereel, gradvr=G))


def test_device_relativistic_assembly_of_a_representation():
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(1.6 / BOHR, 0.3 / BOHR, 2, irrep=5, group=8)
    print(f"relativistic device assembly mmm irrep 5 n={c.nligne}", flush=True)
    _device_vs_host(c, True, False, lambda V, G, E, Ei, Ec, _: c.assemble(V, E, Ei, Ec, True, False, gradvr=G),
                    seed=1)


@pytest.mark.parametrize("branch", [0, 1])
def test_device_relativistic_assembly_with_atoms(branch):
    from fdmnes_amd import cluster as CL
    ga = dict(np.load(ATOMS_G))
    at = (ga["posi"] / BOHR, ga["rmt"] / BOHR, ga["lmax"])
    R, h, lm = float(ga["R"]) / BOHR, float(ga["h"]) / BOHR, int(ga["lmaxso"])
    c, o = CL.Cluster(R, h, lm, atoms=at), O.Cluster(R, h, lm, atoms=at)
    _, _, E_comp, _, basereel = BRANCHES[branch]
    ens = [(0.3 + 0.35 * k, 0.05 if E_comp else 0.0) for k in range(3)]
    phis = [phiato(o, ga["posi"] / BOHR, ga["lmax"], E, Ei, E_comp) for E, Ei in ens]
    print(f"relativistic device assembly with atoms, branch {BRANCHES[branch]} n={c.nligne}", flush=True)
    _device_vs_host(c, E_comp, basereel,
                    lambda V, G, E, Ei, Ec, ph: c.assemble_atoms(V, E, Ei, Ec, E_comp, basereel, ph, gradvr=G),
                    phis=phis, seed=2)


@pytest.mark.parametrize("branch", range(3))
def test_device_relativistic_rows_solve_to_reference(branch):
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import solver as SV
    from fdmnes_amd.assemble import DeviceAssembler
    g = np.load(P1)
    c = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]))
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    dt = torch.complex128 if cal else torch.float64
    a = DeviceAssembler(c, max_points=1, relativistic=True)
    vals = [torch.zeros(c.nnz, dtype=dt, device="cuda")]
    rhs = [torch.zeros((c.nligne, c.nlmso), dtype=dt, device="cuda")]
    a.run([torch.from_numpy(g["V"]).cuda()], [(E, Ei, E)], E_comp, basereel, vals, rhs,
          G=[torch.from_numpy(np.ascontiguousarray(g["gradvr"])).cuda()])
    rp, col, _, _ = c.assemble(g["V"], E, Ei, E, E_comp, basereel, gradvr=g["gradvr"])
    lb1, lb2 = c.profile()
    v, r = vals[0].cpu().numpy(), rhs[0].cpu().numpy()
    ref = g[f"x{branch}"]
    x, _, _ = SV.mat_solve(lb1, lb2, rp, col, v, r, cal)
    xm, sweeps, _ = SV.mat_solve_mixed(lb1, lb2, rp, col, v, r, cal)
    r64 = np.abs(x - ref).max() / np.abs(ref).max()
    r32 = np.abs(xm - ref).max() / np.abs(ref).max()
    print(f"branch {BRANCHES[branch]}: fp64 drop-in vs reference {r64:.2e}, mixed ({sweeps} sweeps) {r32:.2e}",
          flush=True)
    assert r64 <= 1e-10 and r32 <= 1e-10


@pytest.mark.parametrize("branch", range(3))
def test_shim_relativistic_matches_reference_mat_solve(branch):
    from oracle import ref as R
    if not (R.available() and os.path.exists(SHIM)):
        pytest.skip("needs oracle/_ref and the shim (built where the reference sources are)")
    import torch  # noqa: F401  (one HIP runtime for the process)
    from fdmnes_amd import _lib
    _lib.lib()
    R.lib()
    shim = ctypes.CDLL(SHIM, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    g = np.load(P1)
    cl = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]))
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    bes, neu = R.phiso(cl, E_comp, E, Ei, E)
    print(f"relativistic shim: branch {BRANCHES[branch]} n={cl.nligne}", flush=True)
    kw = dict(gradvr=g["gradvr"], cgrad=g["cgrad"])
    x_ref = R.mat_solve(cl, g["V"], E, Ei, bes, neu, E_comp, cal, basereel, **kw)
    x_gpu = R.mat_solve(cl, g["V"], E, Ei, bes, neu, E_comp, cal, basereel, fn=shim.mat_solve_, **kw)
    assert np.array_equal(x_ref, g[f"x{branch}"])
    rel = np.abs(x_gpu - x_ref).max() / np.abs(x_ref).max()
    print("shim vs reference mat_solve", rel, flush=True)
    assert rel <= 1e-10
