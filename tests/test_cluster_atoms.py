"""The product's cluster builder with muffin-tin atoms (fdmgpu_cluster_create_atoms,
fdmgpu_cluster_assemble_atoms) against the reference (tests/golden/
atoms_R1.9_h0.25.npz, three atoms, two of them at the same distance from the
absorber) and the oracle:

- structure: the points inside each sphere (reseau, general.f90:9563-9577),
  the atom border lists and weights (bordure, :9924-10068), Ylmato (Ylmpt,
  :11533-11568), the atom blocks' place among the rows and the profile
  (newind, mat.f90:1790-1884) equal the reference's;
- the compressed rows of the three calcMatRow branches (real, complex
  energy, complex basis) equal the oracle's in structure and within 1e-15
  (relative to the largest entry) in value;
- the product's rows solved by the oracle's elimination reproduce the
  reference's solution at every third row: within 1e-12 in complex
  arithmetic, 1e-10 in the real basis without damping, whose conditioning
  carries the entries' last-bit differences (Ylm) to 3e-11 in the solution."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_atoms import BOHR, BRANCHES, G, phiato


def _clusters():
    from fdmnes_amd import cluster as CL
    g = dict(np.load(G))
    at = (g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"])
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    return g, CL.Cluster(R, h, lm, atoms=at), O.Cluster(R, h, lm, atoms=at)


def test_product_atom_structure_equals_reference():
    g, c, _ = _clusters()
    assert (c.npoint, c.npr, c.nligne) == (g["npoint"], g["npr"], g["nligne"])
    assert np.array_equal(c.numia, g["numia"][: c.npoint])
    for a in range(c.natome):
        d = c.atom(a)
        assert (d["nbord"], d["nbordf"], d["ianew"]) == (g["nbord"][a], g["nbordf"][a], g["ianew"][a])
        assert np.array_equal(d["ibord"] + 1, g[f"ibord{a}"])
        assert np.array_equal(d["poidsa"], g[f"poidsa{a}"])
        assert np.abs(d["ylmato"] - g[f"ylmato{a}"]).max() <= 1e-15
    lb1, lb2 = c.profile()
    assert np.array_equal(lb1, g["lb1"]) and np.array_equal(lb2, g["lb2"])
    # each interstitial point's row is the reference's new(i); atom rows are
    # the gaps newinv marks with -ia
    rows = c.point_rows
    inter = rows >= 0
    assert np.array_equal(rows[inter] + 1, g["new"][1:c.npoint + 1][inter])
    newinv = g["newinv"]  # rows 1..nligne
    for a in range(c.natome):
        d = c.atom(a)
        assert np.all(newinv[d["ianew"]:d["ianew"] + d["nlmsa"]] == -(a + 1))


def test_product_without_atoms_is_unchanged_by_the_atom_path():
    """No atom within the sphere's grid: the same as the plain builder."""
    from fdmnes_amd import cluster as CL
    g = dict(np.load(G))
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    plain = CL.Cluster(R, h, lm)
    with pytest.raises(CL._lib.FdmGpuError):  # an atom the grid does not reach is refused
        CL.Cluster(R, h, lm, atoms=(np.array([[40.0, 0, 0]]), np.array([0.5]), np.array([1])))
    assert plain.natome == 0 and plain.npr == plain.npoint
    assert np.array_equal(plain.point_rows, np.arange(plain.npoint))


@pytest.mark.parametrize("branch", range(3))
def test_product_atom_rows_equal_oracle_and_solve_to_reference(branch):
    g, c, o = _clusters()
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    phi = phiato(o, g["posi"] / BOHR, g["lmax"], E, Ei, E_comp)
    o.set_phiato(phi)
    rp, col, vals, rhs = c.assemble_atoms(g["V"][: c.npoint], E, Ei, E, E_comp, basereel, phi)
    orp, ocol, ovals, orhs = o.assemble_csr(g["V"], E, Ei, E, E_comp, cal, basereel)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(vals - ovals).max() <= 1e-15 * np.abs(ovals).max()
    assert np.abs(rhs - orhs).max() <= 1e-14 * np.abs(orhs).max()
    assert np.iscomplexobj(vals) == cal
    lb1, lb2 = c.profile()
    st, x, fr = O.mat_solve_csr(lb1, lb2, rp, col, vals, rhs, cal)
    assert st == 0, fr
    ref = g[f"x{branch}"]
    rel = np.abs(x[g["rows"]] - ref).max() / np.abs(ref).max()
    print("branch", branch, "product rows, oracle solve vs reference", rel)
    assert rel <= (1e-12 if cal else 1e-10)


def test_complex_radial_solutions_need_complex_arithmetic():
    from fdmnes_amd import cluster as CL
    g, c, o = _clusters()
    phi = phiato(o, g["posi"] / BOHR, g["lmax"], 0.7, 0.05, True)
    with pytest.raises(ValueError):
        c.assemble_atoms(g["V"][: c.npoint], 0.7, 0.0, 0.7, False, True, phi)
    with pytest.raises(CL._lib.FdmGpuError):  # the plain assembler refuses a cluster with atoms
        c.assemble(g["V"][: c.npoint], 0.7, 0.0, 0.7, False, True)


@pytest.mark.parametrize("branch", range(3))
def test_atom_assembly_tables_rebuild_the_host_rows(branch):
    """The device assembler's tables of a cluster with atoms, applied in
    numpy as its kernels apply them (base values, diagonals, outer columns
    and rows with the oracle's outer-sphere functions, the atom couplings and
    atom rows from phiato), give the host assembly's rows: bit for bit where
    no outer-sphere function enters, within 1e-14 elsewhere."""
    g, c, o = _clusters()
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    phi = phiato(o, g["posi"] / BOHR, g["lmax"], E, Ei, E_comp)
    V = g["V"][: c.npoint]
    _, _, hv, hr = c.assemble_atoms(V, E, Ei, E, E_comp, basereel, phi)
    t, at = c.assembly_tables(), c.atom_assembly_tables()
    flat = np.concatenate([p.reshape(-1) for p in phi])
    assert len(flat) == at["nphi"]
    B, N = o.phiso(E_comp, E, Ei)
    cf = N if basereel else B + 1j * N
    nl, nso = c.nlmso, c.nligne - c.nlmso
    lof = np.array([int(np.floor(np.sqrt(k) + 1e-9)) for k in range(nl)])
    Y = t["ylm"].reshape(t["nb"], nl)
    vals = t["base"].astype(np.complex128)
    rhs = np.zeros((c.nligne, nl), np.complex128)
    inside = t["qdiag"] < 0
    assert np.array_equal(inside, c.numia > 0)
    vals[t["qdiag"][~inside]] = V[~inside] - E - t["clapl0"] + (-1j * Ei if E_comp else 0)
    exact = np.ones(c.nnz, bool)
    for r in range(t["nrow_out"]):
        b, w = t["ob"][t["optr"][r]:t["optr"][r + 1]], t["ow"][t["optr"][r]:t["optr"][r + 1]]
        for lm in range(nl):
            vals[t["oq"][r] + lm] = np.sum(cf[b, lof[lm]] * Y[b, lm] * w)
            rhs[t["orow"][r], lm] = -np.sum(Y[b, lm] * B[b, lof[lm]] * w)
        exact[t["oq"][r]:t["oq"][r] + nl] = False
    ns = t["nsortf"]
    for lm in range(nl):
        q = t["row_ptr"][nso + lm] + ns
        for lp in range(nl):
            f = t["wso"] * (Y[:ns, lm] * Y[:ns, lp])
            vals[q + lp] = np.sum(f * cf[:ns, lof[lp]])
            rhs[nso + lm, lp] = -np.sum(t["wso"] * Y[:ns, lm] * Y[:ns, lp] * B[:ns, lof[lp]])
        exact[q:q + nl] = False
    for r in range(at["nar"]):  # grid rows' couplings to an atom's expansion, summed in stencil order
        na = at["nlmsa"][at["aa"][r]]
        s = np.zeros(na, np.complex128)
        for k in range(at["aptr"][r], at["aptr"][r + 1]):
            s = s + at["alw"][k] * flat[at["alphi"][k]:at["alphi"][k] + na]
        vals[at["aq"][r]:at["aq"][r] + na] = s
    k = 0
    po = np.concatenate([[0], np.cumsum(at["nbordf"])])
    yo = np.concatenate([[0], np.cumsum(at["nbordf"] * at["nlmsa"])])
    for a in range(at["natome"]):
        na, nbf = at["nlmsa"][a], at["nbordf"][a]
        pw = at["poidsa"][po[a]:po[a + 1]]
        Ya = at["ylmato"][yo[a]:yo[a + 1]].reshape(nbf, na)
        ph = flat[at["phoff"][a]:at["phoff"][a] + nbf * na].reshape(nbf, na)
        for lm in range(na):
            s = np.zeros(na, np.complex128)
            for b in range(nbf):
                s = s + (pw[b] * (ph[b] * Ya[b, lm]) if cal else pw[b] * Ya[b, lm] * ph[b])
            vals[at["bq"][k]:at["bq"][k] + na] = s
            k += 1
    if not cal:
        vals, rhs = vals.real, rhs.real
    assert np.array_equal(vals[exact], hv[exact])
    assert np.abs(vals - hv).max() <= 1e-14 * np.abs(hv).max()
    assert np.abs(rhs - hr).max() <= 1e-14 * np.abs(hr).max()
