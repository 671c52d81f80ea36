"""The product's cluster builder (host C++) against the oracle: same radial
order, neighbour structure, profile and compressed rows; and the workload
definitions."""
import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067


@pytest.mark.parametrize("R,h,lm", [(1.2, 0.25, 2), (2.0, 0.2, 3), (3.0, 0.24, 2)])
def test_builder_matches_oracle(R, h, lm):
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(R / BOHR, h / BOHR, lm)
    o = O.Cluster(R / BOHR, h / BOHR, lm)
    assert (c.npoint, c.npso, c.nligne, c.nsortf, c.nsort) == (o.npoint, o.npso, o.nligne, o.nsortf, o.nsort)
    lb1, lb2 = c.profile()
    ol1, ol2, _, _ = o.profile()
    assert np.array_equal(lb1, ol1) and np.array_equal(lb2, ol2)
    assert np.array_equal(c.points, o.xyz[: c.npoint, :3])
    V = np.random.default_rng(1).uniform(-1, 0, c.npoint)
    for E_comp, basereel in [(True, False), (False, True), (False, False)]:
        E, Ei = 0.5, (0.03 if E_comp else 0.0)
        cal = E_comp or not basereel
        rp, col, vals, rhs = c.assemble(V, E, Ei, E, E_comp, basereel)
        orp, ocol, ovals, orhs = o.assemble_csr(V, E, Ei, E, E_comp, cal, basereel)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        assert np.abs(vals - ovals).max() <= 1e-14 * np.abs(ovals).max()
        assert np.abs(rhs - orhs).max() <= 1e-13 * np.abs(orhs).max()


def test_workload_sizes():
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    assert len(cfg.lattice_sites("fcc", 3.61, cfg.CU_REAL.atoms_radius)) == 13
    assert len(cfg.lattice_sites("bcc", 2.866, 5.0)) == 59
    c = CL.Cluster(cfg.CU_REAL.R / BOHR, cfg.CU_REAL.h / BOHR, cfg.CU_REAL.lmaxso)
    assert c.npoint == 8217 and c.nligne == 8226


def test_sources_and_potential():
    from fdmnes_amd import configs as cfg
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.5, -0.5]])
    V = cfg.potential(cfg.FE_FP64, pts)
    assert np.all(np.isfinite(V)) and V[0] < -100  # absorber well, regularised at h/2
    s = cfg.sources(cfg.FE_FP64, pts, 5)
    assert s.shape == (5, 16) and np.all(s[3:] == 0)
    assert abs(s[0, 0] - 1 / np.sqrt(4 * np.pi)) < 1e-15
    # the real harmonics agree with the oracle's (reference convention)
    y = cfg.real_ylm(3, pts[1:])
    for p, yy in zip(pts[1:], y):
        assert np.abs(yy - O.ylm_real(3, p, float(np.linalg.norm(p)))).max() < 1e-14


def _from_tables(t, o, V, E, Ei, Ec, E_comp, basereel):
    """One energy point rebuilt from the builder's energy-independent tables
    and the oracle's outer-sphere functions (phiso), in numpy."""
    cal = E_comp or not basereel
    B, N = o.phiso(E_comp, Ec, Ei)  # (nsort, lmax+1)
    cf = N if basereel else B + 1j * N
    nl, npn = t["nlmso"], t["npoint"]
    lof = np.array([int(np.floor(np.sqrt(lm) + 1e-9)) for lm in range(nl)])
    Y = t["ylm"].reshape(t["nb"], nl)
    vals = t["base"].astype(np.complex128)
    rhs = np.zeros((t["nligne"], nl), np.complex128)
    vals[t["qdiag"]] = V - E - t["clapl0"] + (-1j * Ei if E_comp else 0)
    for r in range(t["nrow_out"]):
        k0, k1 = t["optr"][r], t["optr"][r + 1]
        b, w = t["ob"][k0:k1], t["ow"][k0:k1]
        for lm in range(nl):
            vals[t["oq"][r] + lm] = np.sum(cf[b, lof[lm]] * Y[b, lm] * w)
            rhs[t["orow"][r], lm] = -np.sum(Y[b, lm] * B[b, lof[lm]] * w)
    ns = t["nsortf"]
    for lm in range(nl):
        q = t["row_ptr"][npn + lm] + ns
        for lp in range(nl):
            c = t["pw"] * Y[:ns, lm] * Y[:ns, lp]
            vals[q + lp] = np.sum(c * cf[:ns, lof[lp]])
            rhs[npn + lm, lp] = -np.sum(c * B[:ns, lof[lp]])
    return (vals, rhs) if cal else (vals.real, rhs.real)


@pytest.mark.parametrize("R,h,lm", [(1.2, 0.25, 2), (2.0, 0.2, 3)])
def test_assembly_tables_rebuild_oracle_rows(R, h, lm):
    """The tables a device assembler holds, plus the per-energy outer-sphere
    functions, reproduce the oracle's compressed rows and RHS (calcMatRow,
    mat.f90:1142-1511) for the three arithmetic branches."""
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(R / BOHR, h / BOHR, lm)
    o = O.Cluster(R / BOHR, h / BOHR, lm)
    t = c.assembly_tables()
    assert t["nb"] == o.nsort and t["nnz"] == c.nnz
    assert np.array_equal(np.sort(t["orow"]), t["orow"]) and t["optr"][-1] == t["nlink"]
    V = np.random.default_rng(2).uniform(-1, 0, c.npoint)
    for E_comp, basereel in [(True, False), (False, True), (False, False), (True, True)]:
        E, Ei, Ec = 0.7, (0.04 if E_comp else 0.0), 0.9
        cal = E_comp or not basereel
        orp, ocol, ovals, orhs = o.assemble_csr(V, E, Ei, Ec, E_comp, cal, basereel)
        vals, rhs = _from_tables(t, o, V, E, Ei, Ec, E_comp, basereel)
        assert np.array_equal(t["row_ptr"], orp)
        assert np.abs(vals - ovals).max() <= 1e-14 * np.abs(ovals).max()
        assert np.abs(rhs - orhs).max() <= 1e-13 * np.abs(orhs).max()
        # the energy-independent entries are the host assembler's, bit for bit
        fixed = np.ones(c.nnz, bool)
        fixed[t["qdiag"]] = False
        for r in range(t["nrow_out"]):
            fixed[t["oq"][r]: t["oq"][r] + c.nlmso] = False
        for l in range(c.nlmso):
            q = t["row_ptr"][c.npoint + l] + c.nsortf
            fixed[q: q + c.nlmso] = False
        _, _, hvals, _ = c.assemble(V, E, Ei, Ec, E_comp, basereel)
        assert np.array_equal(np.real(hvals)[fixed], t["base"][fixed])
        if cal:
            assert not np.any(np.imag(hvals)[fixed])
