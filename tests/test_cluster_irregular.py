"""The product's builder for the irregular cluster solution
(fdmgpu_cluster_create_atoms_ri, fdmgpu_cluster_assemble_atoms_ri) against
the reference (tests/golden/atoms_tau_R1.9_h0.25.npz: newind and mat_solve
with irregular, n_ato_RI = 2, run from oracle/_ref) and the oracle:

- structure: nligne = npr + 2 sum nlmsa (no outer-sphere rows), each atom's
  two blocks in the radial order, the profile (newind, mat.f90:1761-1921),
  equal the reference's;
- the compressed rows (grid rows coupled to both atomic solutions, the
  incoming wave -(-i j + n) Ylm at outside neighbours, atom rows matched on
  the two halves of the border) equal the oracle's in structure and within
  1e-15 in value;
- solved by the oracle's elimination they give the reference's irregular
  solution within 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_tau_ri import BOHR, _load, _split


def _setup():
    from fdmnes_amd import cluster as CL
    g, t = _load()
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    c = CL.Cluster(R, h, lm, atoms=at, irregular=True)
    o = O.Cluster(R, h, lm, atoms=at, irregular=True)
    phi, phi2 = _split(t["phi_c"], o), _split(t["phi2_c"], o)
    return g, t, c, o, phi, phi2


def test_product_irregular_structure_equals_reference():
    g, t, c, o, _, _ = _setup()
    assert c.nligne == int(t["irr_nligne"]) == c.npr + 2 * sum((int(L) + 1) ** 2 for L in g["lmax"])
    lb1, lb2 = c.profile()
    assert np.array_equal(lb1, t["irr_lb1"]) and np.array_equal(lb2, t["irr_lb2"])
    assert [c.atom(a)["ianew"] for a in range(c.natome)] == list(t["irr_ianew"])
    rows = c.point_rows
    inter = rows >= 0
    assert np.array_equal(rows[inter] + 1, t["irr_new"][1:c.npoint + 1][inter])


def test_product_irregular_rows_equal_oracle_and_solve_to_reference():
    g, t, c, o, phi, phi2 = _setup()
    E, Ei = float(t["Ec"]), float(t["Eimag"])
    o.set_phiato(phi, phi2)
    rp, col, vals, rhs = c.assemble_atoms(t["V"][: c.npoint], E, Ei, E, True, False, phi, phi2)
    orp, ocol, ovals, orhs = o.assemble_csr(t["V"], E, Ei, E, True, True, False)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(vals - ovals).max() <= 1e-15 * np.abs(ovals).max()
    assert np.abs(rhs - orhs).max() <= 1e-14 * np.abs(orhs).max()
    lb1, lb2 = c.profile()
    st, x, fr = O.mat_solve_csr(lb1, lb2, rp, col, vals, rhs, True)
    assert st == 0, fr
    ref = t["x_irr"]
    rel = np.abs(x[t["irr_rows"]] - ref).max() / np.abs(ref).max()
    print("product irregular rows, oracle solve vs reference", rel)
    assert rel <= 1e-12


def test_irregular_assembly_refuses_mismatched_inputs():
    from fdmnes_amd import cluster as CL
    g, t, c, o, phi, phi2 = _setup()
    V, E, Ei = t["V"][: c.npoint], float(t["Ec"]), float(t["Eimag"])
    with pytest.raises(CL._lib.FdmGpuError):  # the irregular rows need both atomic solutions
        c.assemble_atoms(V, E, Ei, E, True, False, phi)
    with pytest.raises(CL._lib.FdmGpuError):  # and complex energy
        c.assemble_atoms(V, E, 0.0, E, False, False, phi, phi2)
    at = c.atom_assembly_tables()  # two row blocks per atom, no outer columns
    assert len(at["bq"]) == 2 * sum(c.atom(a)["nlmsa"] for a in range(c.natome))
    assert np.all(c.assembly_tables()["oq"][:max(c.assembly_tables()["nrow_out"], 1)] == -1)
    reg = CL.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]),
                     atoms=(g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"]))
    with pytest.raises(CL._lib.FdmGpuError):  # a regular cluster takes no irregular solutions
        reg.assemble_atoms(V, E, Ei, E, True, False, phi, phi2)
