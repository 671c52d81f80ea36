"""Symmetry-reduced systems (fdmgpu_cluster_create_mmm and
fdmgpu_cluster_create_group): the representations of the subgroups of mmm —
the mirror groups mmm (eight, wedge x, y, z >= 0), mm2 (four, wedge
x, y >= 0) and m (two, wedge z >= 0), and the groups with rotations or the
inversion, -1 (two), 2 about z (two), 2/m (four) and 222 (four), whose
calculation areas are posequiv's half-spaces and quarter-spaces with their
boundary planes split (general.f90:8950-8978) — each solved on its area,
reproduce the solution of the whole cluster (the oracle's P1 restatement)
for a potential with exactly that symmetry.  With u = rvol * psi on the
wedge (mat.f90:2132-2143), the reduced rows are the whole cluster's rows
restricted to one representation: u / rvol equals the whole solution at the
wedge points and the outer amplitudes are equal, for every right-hand side
(incoming wave) of the representation.  A wrong character, rvol factor,
quadrature weight or harmonic assignment breaks it."""
import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067


MIRRORS = {8: 7, 6: 3, 3: 4}  # point-group number -> mirrors (bit a = normal to axis a)
# point-group number -> its operations as axis sign flips (bit a: coordinate a negated)
OPS = {3: [0, 4], 6: [0, 3, 1, 2], 8: [0, 6, 5, 3, 7, 1, 2, 4], 2: [0, 7], 4: [0, 3], 5: [0, 3, 7, 4],
       7: [0, 6, 5, 3]}
GROUPS = [3, 6, 8, 2, 4, 5, 7]


def _chars(group, v):
    return tuple(bin(g & v).count("1") % 2 for g in OPS[group])


def _irreps(group):
    """The representations, each named by the smallest mask v whose
    characters (-1)^popcount(g & v) it has."""
    out, seen = [], set()
    for v in range(8):
        c = _chars(group, v)
        if c not in seen:
            seen.add(c)
            out.append(v)
    assert len(out) == len(OPS[group])
    return out


def _flip(pts, g):
    return pts * np.array([(-1) ** ((g >> a) & 1) for a in range(3)])


def _potential(pts, group=8):
    """Invariant under the group's operations and no others: a function with
    no symmetry summed over the group."""
    def f(q):
        x, y, z = q[:, 0], q[:, 1], q[:, 2]
        r2 = x * x + y * y + z * z
        v = -1.3 * np.exp(-r2 / 6.0) + 0.3 * np.cos(0.9 * x) * np.cos(0.6 * y) * np.cos(1.3 * z) - 0.2
        return v + 0.25 * np.exp(-r2 / 8.0) * (np.tanh(0.5 * x + 0.2) + 0.6 * np.sin(0.4 * y + 0.3 * x)
                                               + np.tanh(0.7 * z - 0.1 * y))
    return sum(f(_flip(pts, g)) for g in OPS[group]) / len(OPS[group])


def _rep(CL, R, h, lm, g, group=8):
    """The representation's system, or None when no harmonic up to lm has
    its parities (mmm's irrep 7, odd in x, y and z, needs l >= 3)."""
    from fdmnes_amd import _lib
    try:
        return CL.Cluster(R, h, lm, irrep=g, group=group)
    except _lib.FdmGpuError:
        assert lm < 3 and g == 7 and group == 8
        return None


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("group", GROUPS)
def test_harmonic_parities(group):
    """Every real harmonic of a representation transforms under each of the
    group's operations with the representation's character (checked on the
    harmonics themselves), and the representations share out every harmonic
    once."""
    from fdmnes_amd import cluster as CL
    lmax = 4
    v = np.array([0.31, 0.52, 0.79])
    y0 = O.ylm_real(lmax, v, float(np.linalg.norm(v)))
    seen = []
    for g in _irreps(group):
        c = CL.Cluster(1.2 / BOHR, 0.25 / BOHR, lmax, irrep=g, group=group)
        for lm in c.harmonics:
            for op in OPS[group]:
                w = _flip(v[None, :], op)[0]
                ya = O.ylm_real(lmax, w, float(np.linalg.norm(w)))[lm]
                chi = (-1) ** (bin(op & g).count("1") % 2)
                assert abs(ya - chi * y0[lm]) <= 1e-12 * max(1.0, abs(y0[lm])), (g, lm, op)
        seen += list(c.harmonics)
    assert sorted(seen) == list(range((lmax + 1) ** 2))


@pytest.mark.parametrize("group", GROUPS)
def test_wedge_grid_and_weights(group):
    from fdmnes_amd import cluster as CL
    full = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2)
    c = CL.Cluster(1.6 / BOHR, 0.25 / BOHR, 2, irrep=0, group=group)
    pts, fp = c.points, full.points
    if group in MIRRORS:
        for a in range(3):
            if (MIRRORS[group] >> a) & 1:
                assert (pts[:, a] >= 0).all()
            else:
                assert (pts[:, a] < 0).any()
    order = len(OPS[group])
    signs = [[(-1) ** ((f >> a) & 1) for a in range(3)] for f in OPS[group]]
    # every interior point of the whole cluster is the image of exactly one wedge point
    imgs = set()
    for p, rv in zip(pts, c.rvol):
        orbit = {tuple(s * p) for s in np.array(signs)}
        assert rv == len(orbit) / order
        new = {tuple(np.round(np.array(q) / full.adimp).astype(int)) for q in orbit}
        assert not (imgs & new)
        imgs |= new
    assert imgs == {tuple(q) for q in np.round(fp / full.adimp).astype(int)}
    # rows in radial order, as reseau's exchange sort leaves them
    r = np.linalg.norm(pts, axis=1)
    assert (np.diff(r) > -1e-4 * full.adimp).all()


@pytest.mark.parametrize("group", GROUPS)
@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_representations_reproduce_whole_cluster(E_comp, basereel, group):
    from fdmnes_amd import cluster as CL
    R, h, lm = 1.6 / BOHR, 0.25 / BOHR, 2
    o = O.Cluster(R, h, lm)
    cal = E_comp or not basereel
    E, Ei, Ec = 0.35, (0.0367 if E_comp else 0.0), 0.55
    xyz = o.xyz[:o.npoint, :3]
    st, xf, *_ = o.solve(_potential(xyz, group), E, Ei, Ec, E_comp, cal, basereel)
    assert st == 0
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    nsof = o.nligne - o.nlmso
    scale = np.abs(xf).max()
    bound = 1e-9 if basereel else 1e-10
    covered = []
    for g in _irreps(group):
        c = _rep(CL, R, h, lm, g, group)
        if c is None:
            continue
        pts = c.points
        lb1, lb2 = c.profile()
        rp, col, vals, rhs = c.assemble(_potential(pts, group), E, Ei, Ec, E_comp, basereel)
        st, u, _ = O.mat_solve_csr(lb1, lb2, rp, col, vals, rhs, cal)
        assert st == 0
        hm = c.harmonics
        covered += list(hm)
        rows = [row[tuple(k)] for k in np.round(pts / h).astype(int)]
        psi = u[:c.npoint] / c.rvol[:, None]
        d_grid = np.abs(psi - xf[rows][:, hm]).max() / scale
        d_out = np.abs(u[c.npoint:] - xf[nsof + hm][:, hm]).max() / scale
        print("group", group, "irrep", g, "harmonics", list(hm), "rows", c.nligne, "grid", d_grid, "outer", d_out)
        assert d_grid <= bound and d_out <= bound
    assert sorted(covered) == list(range(o.nlmso))
    # the whole solution has no outer amplitude across representations
    gof = {}
    for g in _irreps(group):
        c = _rep(CL, R, h, lm, g, group)
        for k in ([] if c is None else c.harmonics):
            gof[int(k)] = g
    for L in range(o.nlmso):
        for Lp in range(o.nlmso):
            if gof[L] != gof[Lp]:
                assert abs(xf[nsof + Lp, L]) <= bound * scale


@pytest.mark.parametrize("group", GROUPS)
@pytest.mark.parametrize("E_comp,basereel", [(True, False), (False, True)])
def test_assembly_tables_rebuild_representation_rows(E_comp, basereel, group):
    """The device assembler's tables of each representation (stencil weights
    with characters and rvol, per-point outer weights, mirror terms of the
    diagonal, l of each outer unknown) rebuild the host assembler's rows:
    energy-independent entries and diagonals bit for bit, the outer-sphere
    entries and RHS to the last bits of the Bessel/Neumann values."""
    from fdmnes_amd import cluster as CL
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    cal = E_comp or not basereel
    E, Ei, Ec = 0.45, (0.0367 if E_comp else 0.0), 0.7
    for g in _irreps(group):
        c = CL.Cluster(R, h, lm, irrep=g, group=group)
        t = c.assembly_tables()
        V = _potential(c.points, group)
        _, _, hv, hr = c.assemble(V, E, Ei, Ec, E_comp, basereel)
        nl, npn, ns = t["nlmso"], t["npoint"], t["nsortf"]
        fixed = np.ones(c.nnz, bool)
        fixed[t["qdiag"]] = False
        for r in range(t["nrow_out"]):
            fixed[t["oq"][r]: t["oq"][r] + nl] = False
        for k in range(nl):
            q = t["row_ptr"][npn + k] + ns
            fixed[q: q + nl] = False
        assert np.array_equal(np.real(hv)[fixed], t["base"][fixed])
        d = V - E - t["clapl0"]
        ds = t["dself"].reshape(npn, 3)
        for a in range(3):
            d = d + ds[:, a]
        assert np.array_equal(np.real(hv)[t["qdiag"]], d)
        # outer-sphere entries from the tables
        k = np.sqrt(complex(Ec, Ei)) if E_comp else np.sqrt(Ec)
        fn = np.sqrt(k / np.pi)
        Bv = np.zeros((t["nb"], lm + 1), complex)
        Nv = np.zeros((t["nb"], lm + 1), complex)
        for b in range(t["nb"]):
            if E_comp:
                Bv[b], Nv[b] = O.cbessneu(fn, k * t["brad"][b], lm)
            else:
                Bv[b], Nv[b] = O.cbessneur(fn, k * t["brad"][b], lm)
        cf = Nv if basereel else Bv + 1j * Nv
        Y = t["ylm"].reshape(t["nb"], nl)
        L = t["lharm"]
        vals = np.array(hv, complex)
        rhs = np.zeros((t["nligne"], nl), complex)
        for r in range(t["nrow_out"]):
            k0, k1 = t["optr"][r], t["optr"][r + 1]
            b, w = t["ob"][k0:k1], t["ow"][k0:k1]
            for m in range(nl):
                vals[t["oq"][r] + m] = np.sum(cf[b, L[m]] * Y[b, m] * w)
                rhs[t["orow"][r], m] = -np.sum(Y[b, m] * Bv[b, L[m]] * w)
        for m in range(nl):
            q = t["row_ptr"][npn + m] + ns
            for p in range(nl):
                cc = t["wso"] * Y[:ns, m] * Y[:ns, p]
                vals[q + p] = np.sum(cc * cf[:ns, L[p]])
                rhs[npn + m, p] = -np.sum(cc * Bv[:ns, L[p]])
        assert np.abs(vals - hv).max() <= 1e-13 * np.abs(hv).max()
        assert np.abs(rhs - hr).max() <= 1e-13 * np.abs(hr).max()
    assert (t["dself"] != 0).any()


def test_group_and_irrep_checks():
    """A representation of a subgroup of mmm is named by the smallest mask
    of its characters; the groups beyond 4/m (number 11) are not built."""
    from fdmnes_amd import _lib
    from fdmnes_amd import cluster as CL
    R, h = 1.2 / BOHR, 0.25 / BOHR
    # a representation's mask must be the smallest with its characters; no group 12
    for group, irrep in [(6, 4), (3, 1), (3, 2), (5, 2), (2, 2), (4, 4), (7, 4), (8, 8), (12, 0)]:
        with pytest.raises(_lib.FdmGpuError):
            CL.Cluster(R, h, 2, irrep=irrep, group=group)
    whole = CL.Cluster(R, h, 2)
    for group in GROUPS:  # irrep -1: the whole cluster whatever the group
        c = CL.Cluster(R, h, 2, irrep=-1, group=group)
        assert (c.nligne, c.nnz) == (whole.nligne, whole.nnz)
        assert np.array_equal(c.points, whole.points)
