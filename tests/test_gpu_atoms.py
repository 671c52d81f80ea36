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
