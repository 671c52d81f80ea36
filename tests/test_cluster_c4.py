"""Complex harmonics in the outer sphere (the reference's Ylm_comp) and the
four-fold point groups 4, -4 and 4/m (numbers 9, 10, 11), whose
one-dimensional representations have complex characters (Kar/Kari;
calcMatRow's Repres_comp branch, mat.f90:1164-1223; Yc, mat.f90:1733):

- the oracle's Ylm_comp rows against the reference's mat_solve
  (tests/golden/ylmc_R1.2_h0.25.npz): bit for bit at real energy, within
  1e-13 at complex energy;
- the product's P1 rows with complex harmonics equal the oracle's;
- each Yc(l, m) the builder assigns to a representation transforms under
  every operation g as Y(g r) = chi(g) Y(r), evaluated on random
  directions with the oracle's real harmonics (a wrong phase or a swapped
  +-m fails), and every harmonic is assigned once over the representations;
- every point of the whole grid is the image of exactly one area point,
  with rvol = its orbit size / the group's order;
- each representation, solved on its calculation area, reproduces the
  oracle's whole-cluster solution in complex harmonics for a potential with
  exactly the group's symmetry (u / rvol on the area points, the outer
  amplitudes), for 4, -4, 4/m and, with complex harmonics, the subgroups of
  2/m."""
import os

import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067
HERE = os.path.dirname(os.path.abspath(__file__))
YLMC = os.path.join(HERE, "golden", "ylmc_R1.2_h0.25.npz")
YLM_BRANCHES = [(0.7, 0.05, True, False), (0.7, 0.0, False, False), (0.7, 0.0, False, True)]

# the groups' operations as the reference's matrices (opsym, general.f90:4587;
# tests/golden/grid_g*.npz hold the reference's own), p -> M p
_E = np.eye(3)
_C4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
_MZ = np.diag([1, 1, -1])
OPM = {9: [_E, _C4, _C4.T, np.diag([-1, -1, 1])],
       10: [_E, np.diag([-1, -1, 1]), _MZ @ _C4, _MZ @ _C4.T],
       11: [_E, _C4, _C4.T, np.diag([-1, -1, 1]), -_E, _MZ @ _C4, _MZ @ _C4.T, _MZ],
       2: [_E, -_E], 3: [_E, _MZ], 4: [_E, np.diag([-1, -1, 1])],
       5: [_E, np.diag([-1, -1, 1]), -_E, _MZ]}
IRREPS = {9: range(4), 10: range(4), 11: range(8), 2: (0, 1), 3: (0, 4), 4: (0, 1), 5: (0, 1, 4, 5)}


def chi(group, M, v):
    """The builder's character of operation M in representation v: for 4,
    -4, 4/m exp(-i k q pi/2) for a rotation by q pi/2 about z, times p for a
    reflection of z (v = k + 4 [p = -1]); for the mask groups
    (-1)^(number of negated axes in v)."""
    if group in (9, 10, 11):
        q = int(np.rint(np.arctan2(M[1, 0], M[0, 0]) / (np.pi / 2))) % 4
        c = (-1j) ** ((v & 3) * q)
        if M[2, 2] < 0 and (v & 4):
            c = -c
        return c
    flips = [a for a in range(3) if M[a, a] < 0]
    return (-1.0) ** sum((v >> a) & 1 for a in flips)


def yc(l, m, v):
    """Yc(l, m) (mat.f90:1733) of direction v from the oracle's real harmonics."""
    y = O.ylm_real(l, v, float(np.linalg.norm(v)))
    y1, y2 = y[l * l + l + m], y[l * l + l - m]
    if m == 0:
        return complex(y1)
    if m < 0:
        return complex(y2, -y1) / np.sqrt(2)
    return complex(y1, y2) * (-1) ** m / np.sqrt(2)


def potential(pts, group):
    """A potential with exactly the group's symmetry: a function without
    symmetry averaged over the group."""
    def f(q):
        x, y, z = q[:, 0], q[:, 1], q[:, 2]
        r2 = x * x + y * y + z * z
        return -1.3 * np.exp(-r2 / 6.0) + 0.25 * np.exp(-r2 / 8.0) * (
            np.tanh(0.5 * x + 0.2) + 0.6 * np.sin(0.4 * y + 0.3 * x) + np.tanh(0.7 * z - 0.1 * y))
    return sum(f(pts @ M.T) for M in OPM[group]) / len(OPM[group])


@pytest.mark.parametrize("branch", range(3))
def test_oracle_ylm_comp_matches_reference(branch):
    g = np.load(YLMC)
    o = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]))
    o.set_ylm_comp(True)
    E, Ei, E_comp, basereel = YLM_BRANCHES[branch]
    st, x, *_ = o.solve(g["V"], E, Ei, E, E_comp, True, basereel)
    assert st == 0
    ref = g[f"x{branch}"]
    rel = np.abs(x - ref).max() / np.abs(ref).max()
    assert rel <= (1e-13 if E_comp else 0.0), rel


@pytest.mark.parametrize("branch", range(3))
def test_product_ylm_comp_rows_equal_oracle(branch):
    from fdmnes_amd import cluster as CL
    g = np.load(YLMC)
    R, h, lm = float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"])
    o = O.Cluster(R, h, lm)
    o.set_ylm_comp(True)
    c = CL.Cluster(R, h, lm, ylm_comp=True)
    E, Ei, E_comp, basereel = YLM_BRANCHES[branch]
    rp, col, vals, rhs = c.assemble(g["V"], E, Ei, E, E_comp, basereel)
    orp, ocol, ovals, orhs = o.assemble_csr(g["V"], E, Ei, E, E_comp, True, basereel)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.iscomplexobj(vals)
    assert np.abs(vals - ovals).max() <= 1e-15 * np.abs(ovals).max()
    assert np.abs(rhs - orhs).max() <= 1e-14 * np.abs(orhs).max()
    lb1, lb2 = c.profile()
    st, x, _ = O.mat_solve_csr(lb1, lb2, rp, col, vals, rhs, True)
    assert st == 0
    assert np.abs(x - g[f"x{branch}"]).max() <= 1e-12 * np.abs(g[f"x{branch}"]).max()


@pytest.mark.parametrize("group", [9, 10, 11])
def test_harmonics_transform_with_their_characters(group):
    from fdmnes_amd import cluster as CL
    lm = 4
    rng = np.random.default_rng(group)
    dirs = rng.normal(size=(3, 3))
    seen = []
    for v in IRREPS[group]:
        c = CL.Cluster(1.2 / BOHR, 0.3 / BOHR, lm, irrep=v, group=group)
        for k in c.harmonics:
            l = int(np.floor(np.sqrt(k) + 1e-9))
            m = int(k) - l * l - l
            for M in OPM[group]:
                for r in dirs:
                    assert abs(yc(l, m, M @ r) - chi(group, M, v) * yc(l, m, r)) <= 1e-12, (v, l, m)
        seen += list(c.harmonics)
    assert sorted(seen) == list(range((lm + 1) ** 2))


@pytest.mark.parametrize("group", [9, 10, 11])
def test_area_covers_the_grid_once(group):
    from fdmnes_amd import cluster as CL
    full = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    c = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2, irrep=0, group=group)
    h = full.adimp
    order = len(OPM[group])
    imgs = set()
    for p, rv in zip(np.rint(c.points / h).astype(int), c.rvol):
        orbit = {tuple(int(t) for t in np.rint(M @ p)) for M in OPM[group]}
        assert rv == len(orbit) / order
        assert not (imgs & orbit)
        imgs |= orbit
    assert imgs == {tuple(q) for q in np.rint(full.points / h).astype(int)}


@pytest.mark.parametrize("group", [9, 10, 11, 2, 3, 4, 5])
def test_complex_harmonic_representations_reproduce_whole_cluster(group):
    from fdmnes_amd import cluster as CL
    R, h, lm = 1.6 / BOHR, 0.25 / BOHR, 3
    o = O.Cluster(R, h, lm)
    o.set_ylm_comp(True)
    xyz = o.xyz[:o.npoint, :3]
    E, Ei, Ec = 0.35, 0.0367, 0.55
    st, xf, *_ = o.solve(potential(xyz, group), E, Ei, Ec, True, True, False)
    assert st == 0
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    nsof = o.nligne - o.nlmso
    scale = np.abs(xf).max()
    covered = []
    for v in IRREPS[group]:
        c = CL.Cluster(R, h, lm, irrep=v, group=group, ylm_comp=True)
        rp, col, vals, rhs = c.assemble(potential(c.points, group), E, Ei, Ec, True, False)
        lb1, lb2 = c.profile()
        st, u, _ = O.mat_solve_csr(lb1, lb2, rp, col, vals, rhs, True)
        assert st == 0
        hm = c.harmonics
        covered += list(hm)
        rows = [row[tuple(k)] for k in np.round(c.points / h).astype(int)]
        d_grid = np.abs(u[:c.npoint] / c.rvol[:, None] - xf[rows][:, hm]).max() / scale
        d_out = np.abs(u[c.npoint:] - xf[nsof + hm][:, hm]).max() / scale
        print("group", group, "irrep", v, "rows", c.nligne, "harmonics", list(hm), "grid", d_grid, "outer", d_out)
        assert d_grid <= 1e-10 and d_out <= 1e-10
    assert sorted(covered) == list(range(o.nlmso))


def test_complex_groups_need_complex_harmonics():
    from fdmnes_amd import _lib
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(1.2 / BOHR, 0.3 / BOHR, 2, irrep=1, group=9)
    assert c.ylm_comp
    with pytest.raises(_lib.FdmGpuError):  # mm2's mirrors mix Yc(l, m) and Yc(l, -m)
        CL.Cluster(1.2 / BOHR, 0.3 / BOHR, 2, irrep=0, group=6, ylm_comp=True)
    with pytest.raises(_lib.FdmGpuError):  # k = 0..3 for 4
        CL.Cluster(1.2 / BOHR, 0.3 / BOHR, 2, irrep=4, group=9)


def _emulate_complex(c, V, E, Ei, Ec, E_comp, basereel):
    """The device assembler's kernels, in numpy, on the complex tables
    (fdmgpu_cluster_assembly_tables_complex), with the outer functions from
    scipy: rows and RHS of one energy."""
    from scipy.special import spherical_jn, spherical_yn
    t = c.assembly_tables()
    nl, nso = c.nlmso, c.nligne - c.nlmso
    lof = t["lharm"]
    Y = (t["ylm"] + 1j * t["ylm_im"]).reshape(t["nb"], nl)
    k = np.sqrt(Ec + 1j * Ei) if E_comp else np.sqrt(Ec)
    fn = np.sqrt(k / np.pi)
    z = k * t["brad"]
    B = np.stack([fn * spherical_jn(l, z) for l in range(c.lmaxso + 1)], axis=1)
    N = np.stack([fn * spherical_yn(l, z) for l in range(c.lmaxso + 1)], axis=1)
    cf = N if basereel else B + 1j * N
    vals = t["base"] + 1j * t["base_im"]
    rhs = np.zeros((c.nligne, nl), np.complex128)
    qd = t["qdiag"]
    ds = (t["dself"] + 1j * t["dself_im"]).reshape(-1, 3).sum(axis=1)
    vals[qd] = (V - E - t["clapl0"]) + (-1j * Ei if E_comp else 0) + ds
    w_all = t["ow"] + 1j * t["ow_im"]
    for r in range(t["nrow_out"]):
        k0, k1 = t["optr"][r], t["optr"][r + 1]
        b, w = t["ob"][k0:k1], w_all[k0:k1]
        for lm in range(nl):
            vals[t["oq"][r] + lm] = np.sum(cf[b, lof[lm]] * Y[b, lm] * w)
            rhs[t["orow"][r], lm] = -np.sum(B[b, lof[lm]] * Y[b, lm] * w)
    ns = t["nsortf"]
    for lm in range(nl):
        q = t["row_ptr"][nso + lm] + ns
        for lp in range(nl):
            f = t["wso"] * (np.conj(Y[:ns, lm]) * Y[:ns, lp])
            vals[q + lp] = np.sum(f * cf[:ns, lof[lp]])
            rhs[nso + lm, lp] = -np.sum(f * B[:ns, lof[lp]])
    return vals, rhs


@pytest.mark.parametrize("group,irrep", [(9, 1), (9, 2), (10, 3), (11, 5), (1, -1), (5, 1)])
@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_complex_assembly_tables_rebuild_the_host_rows(group, irrep, E_comp, basereel):
    """The device assembler's complex tables (imaginary parts of the
    couplings, outer links and harmonics; complex own-image terms) rebuild
    the host's rows: energy-independent entries exactly, the rest to the
    outer functions' accuracy (scipy here, the recurrence on the host)."""
    from fdmnes_amd import cluster as CL
    kw = dict(irrep=irrep, group=group) if irrep >= 0 else dict(ylm_comp=True)
    if group == 5:  # 2/m: complex harmonics on request (its operations keep Yc(l, m) apart from Yc(l, -m))
        kw["ylm_comp"] = True
    c = CL.Cluster(1.3 / BOHR, 0.3 / BOHR, 2, **kw)
    assert c.ylm_comp
    rng = np.random.default_rng(group * 10 + irrep)
    V = rng.uniform(-1.5, -0.5, c.npoint)
    E, Ei, Ec = 0.9, (0.04 if E_comp else 0.0), 0.8
    rp, col, hv, hr = c.assemble(V, E, Ei, Ec, E_comp, basereel)
    ev, er = _emulate_complex(c, V, E, Ei, Ec, E_comp, basereel)
    t = c.assembly_tables()
    fixed = np.ones(c.nnz, bool)
    fixed[t["qdiag"]] = False
    for r in range(t["nrow_out"]):
        fixed[t["oq"][r]:t["oq"][r] + c.nlmso] = False
    nso = c.nligne - c.nlmso
    for lm in range(c.nlmso):
        q = t["row_ptr"][nso + lm] + t["nsortf"]
        fixed[q:q + c.nlmso] = False
    assert np.array_equal(ev[fixed], hv[fixed])
    assert np.abs(ev[t["qdiag"]] - hv[t["qdiag"]]).max() <= 1e-14 * np.abs(hv).max()
    sc = np.abs(hv).max()
    assert np.abs(ev - hv).max() <= 1e-10 * sc
    assert np.abs(er - hr).max() <= 1e-10 * np.abs(hr).max()
    assert np.abs(hv.imag).max() > 0  # the systems are complex
