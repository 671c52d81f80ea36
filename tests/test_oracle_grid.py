"""Grid builders against the reference's own: point order, neighbour table,
stencil weights and outer-sphere boundary lists and weights from reseau,
laplac and bordure (general.f90:9439, 9732, 9895, run from oracle/_ref by
tests/golden/make_golden.py), for the oracle's restatement and the product's
builder, point group 1, the subgroups of mmm and the four-fold groups 4, -4,
4/m.

Beside the reference, checks that the mathematics fixes, each of which a
plausible mistake fails:
- the stencil reproduces the Laplacian of every monomial of degree <= 5 at
  every point whose twelve neighbours exist (a swapped pair of weights or a
  neighbour one step off fails);
- every neighbour sits at its stencil offset; the points are every lattice
  point within the build radius, once each, in non-decreasing radius;
- the boundary lists are the interior points with an outside neighbour along
  the axes and the outside points with an interior neighbour, and their
  weights add up to 4 pi;
- for the groups, the product's grid-row couplings are the reference's
  -clapl(v) chi(op) rvol(i) / rvol(j) (mat.f90:1159-1233) with the neighbour
  and its symmetry operation from laplac and chi the representation's
  (complex, for 4, -4, 4/m) character of that operation (between points
  whose wave function the representation does not force to zero)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P1 = ["grid_p1_R1.2_h0.25", "grid_p1_R1.6_h0.3", "grid_p1_R2.0_h0.25"]
GROUPS = ["grid_g3_R1.6_h0.25", "grid_g6_R1.6_h0.25", "grid_g8_R1.6_h0.25", "grid_g2_R1.6_h0.25",
          "grid_g4_R1.6_h0.25", "grid_g5_R1.6_h0.25", "grid_g7_R1.6_h0.25", "grid_g9_R1.6_h0.25",
          "grid_g10_R1.6_h0.25", "grid_g11_R1.6_h0.25"]
EPSPOS = 1e-4
# stencil offsets in laplac's order (general.f90:9768-9781): per order io = 1, 2,
# per axis x, y, z, the negative then the positive step
OFFSETS = np.array([[s * io * (a == k) for k in range(3)] for io in (1, 2) for a in range(3) for s in (-1, 1)])


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


@pytest.mark.parametrize("name", P1)
def test_oracle_grid_equals_reference(name):
    g = _load(name)
    c = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, 2)
    assert (c.npoint, c.npso, c.nsortf, c.nsort) == (g["npoint"], g["npso"], g["nsortf"], g["nsort"])
    assert np.array_equal(c.indice, g["indice"])
    assert np.array_equal(c.xyz[:, 3], g["radius"])
    assert np.array_equal(c.ivois - 1, g["ivois"])
    assert np.array_equal(c.clapl, g["clapl"])
    isrt, poidso = c.boundary
    assert np.array_equal(isrt - 1, g["isrt"])
    assert np.array_equal(poidso, g["poidso"])


@pytest.mark.parametrize("name", P1)
def test_product_grid_equals_reference(name):
    from fdmnes_amd import cluster as CL
    g = _load(name)
    h = float(g["h"]) / BOHR
    c = CL.Cluster(float(g["R"]) / BOHR, h, 2)
    assert (c.npoint, c.npso, c.nsortf, c.nsort) == (g["npoint"], g["npso"], g["nsortf"], g["nsort"])
    assert np.array_equal(np.rint(c.points / h).astype(int), g["indice"][: g["npoint"]])
    t = c.assembly_tables()
    assert np.array_equal(t["wso"], g["poidso"])
    assert np.array_equal(t["brad"], g["radius"][g["isrt"]])
    assert t["clapl0"] == g["clapl"][0]


def _opmat(g, op):
    """The reference's matrix of operation `op` (the fixture's copy of opsym)."""
    return g["opmat"][list(g["ops"]).index(abs(int(op)))]


def _chi(g, op, irrep):
    """The builder's character of the reference's operation `op` in the
    representation `irrep`."""
    from .test_cluster_c4 import chi
    return chi(int(g["group"]), _opmat(g, op), irrep)


@pytest.mark.parametrize("name", GROUPS)
def test_product_mirror_group_builder_equals_reference(name):
    from fdmnes_amd import cluster as CL
    g = _load(name)
    grp, h = int(g["group"]), float(g["h"]) / BOHR
    npt = int(g["npoint"])
    from .test_cluster_c4 import IRREPS
    from .test_cluster_mmm import _irreps
    for irrep in (IRREPS[grp] if grp in (9, 10, 11) else _irreps(grp)):
        c = CL.Cluster(float(g["R"]) / BOHR, h, 3, irrep=irrep, group=grp)
        assert (c.npoint, c.npso, c.nsortf, c.nsort) == (npt, g["npso"], g["nsortf"], g["nsort"])
        assert np.array_equal(np.rint(c.points / h).astype(int), g["indice"][:npt])
        assert np.array_equal(c.rvol, g["rvol"])
        assert np.array_equal(c.outer_weights(), g["poidso"])
        # grid-row couplings between interior points, energy-independent
        rp, col, vals, _ = c.assemble(np.zeros(npt), 0.0, 0.0, 1.0, False, True)
        rvol, clapl = g["rvol"], g["clapl"]
        # points an operation of odd character leaves in place (a mirror
        # plane, a two-fold axis, the centre of inversion) carry a zero wave
        # function: there two operations fold a neighbour onto the same image
        # with different characters (the reference takes the first in its
        # list, e.g. C2z where the product takes a mirror), and the couplings
        # among such points are left out of the comparison
        ind = g["indice"]
        zero = np.zeros(len(ind), bool)
        for op in g["ops"]:  # points the operation leaves in place, with a character other than 1
            if abs(_chi(g, op, irrep) - 1) > 1e-12:
                zero |= np.all(ind @ _opmat(g, op).T == ind, axis=1)
        worst = 0.0
        for i in range(npt):
            if zero[i]:
                continue
            want = {}
            for v in range(12):
                j, op = int(g["ivois"][i, v]), int(g["isvois"][i, v])
                if j < npt and j != i and not zero[j]:
                    want[j] = want.get(j, 0.0) - clapl[v + 1] * _chi(g, op, irrep) * rvol[i] / rvol[j]
            got = {int(k) - 1: complex(x) for k, x in zip(col[rp[i]:rp[i + 1]], vals[rp[i]:rp[i + 1]])
                   if int(k) - 1 < npt and int(k) - 1 != i and not zero[int(k) - 1]}
            for j in set(want) | set(got):
                worst = max(worst, abs(want.get(j, 0.0) - got.get(j, 0.0)) / abs(clapl[1]))
        assert worst <= 1e-14, (grp, irrep, worst)


def _laplacian_residual(g, clapl, ivois, xyz):
    """max over monomials x^a y^b z^c (a + b + c <= 5) and complete-stencil
    points of |sum_v clapl(v) f(x_j) + clapl(0) f(x_i) - lap f(x_i)|, relative
    to the largest stencil term."""
    full = np.all(ivois >= 0, axis=1)
    i = np.nonzero(full)[0]
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    worst = 0.0
    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                f = x ** a * y ** b * z ** c
                lap = (a * (a - 1) * x ** max(a - 2, 0) * y ** b * z ** c + b * (b - 1) * x ** a * y ** max(b - 2, 0)
                       * z ** c + c * (c - 1) * x ** a * y ** b * z ** max(c - 2, 0))
                s = clapl[0] * f[i] + np.sum(clapl[1:][None, :] * f[ivois[i]], axis=1)
                scale = np.abs(clapl).max() * np.abs(f).max() + 1e-300
                worst = max(worst, float(np.abs(s - lap[i]).max() / scale))
    return worst


@pytest.mark.parametrize("name", P1)
def test_stencil_is_exact_on_quintics(name):
    g = _load(name)
    h = float(g["h"]) / BOHR
    xyz = g["indice"] * h
    assert _laplacian_residual(g, g["clapl"], g["ivois"], xyz) <= 1e-12
    # planted mistakes: the first- and second-neighbour weights of x swapped,
    # and the +x neighbour replaced by the +2x one
    bad = g["clapl"].copy()
    bad[[2, 8]] = bad[[8, 2]]
    assert _laplacian_residual(g, bad, g["ivois"], xyz) > 1e-6
    badv = g["ivois"].copy()
    badv[:, 1] = g["ivois"][:, 7]
    assert _laplacian_residual(g, g["clapl"], badv, xyz) > 1e-6
    # the weights themselves: 4/(3 h^2) and -1/(12 h^2), clapl(0) = -sum
    assert np.allclose(g["clapl"][1:7], 4 / (3 * h * h), rtol=1e-15)
    assert np.allclose(g["clapl"][7:], -1 / (12 * h * h), rtol=1e-15)
    assert g["clapl"][0] == -np.sum(g["clapl"][1:])


@pytest.mark.parametrize("name", P1 + GROUPS)
def test_neighbours_points_and_boundary(name):
    g = _load(name)
    R, h = float(g["R"]) / BOHR, float(g["h"]) / BOHR
    ind, rad, npt = g["indice"], g["radius"], int(g["npoint"])
    n_iv = g["ivois"].shape[0]
    # neighbours at their offsets (the point itself or its image for a group)
    for v in range(12):
        j = g["ivois"][:n_iv, v]
        ok = j >= 0
        if "group" in g:
            # the neighbour is the image of the offset point under the
            # inverse of the operation laplac records for it (posequiv
            # records the operation that maps the area point back)
            ops = g["isvois"][:, v][ok]
            want = ind[:n_iv][ok] + OFFSETS[v]
            for k, (jj, w) in enumerate(zip(j[ok], want)):
                if k < len(ops):
                    assert np.array_equal(ind[jj], np.rint(_opmat(g, ops[k]).T @ w).astype(int))
        else:
            assert np.array_equal(ind[j[ok]], ind[:n_iv][ok] + OFFSETS[v])
    # radial order and radii
    assert np.all(np.diff(rad) > -EPSPOS)
    assert np.allclose(rad, h * np.linalg.norm(ind, axis=1), rtol=0, atol=1e-12)
    # every lattice point within the build radius once (P1), or one per orbit (groups)
    rmax = R / h + np.sqrt(2.0) * 4 / 2 + EPSPOS
    m = int(np.ceil(rmax)) + 1
    r = np.arange(-m, m + 1)
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    inside = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) <= rmax
    lattice = {tuple(p) for p in np.stack([X[inside], Y[inside], Z[inside]], 1)}
    got = [tuple(p) for p in ind]
    assert len(set(got)) == len(got)
    if "group" in g:
        # one point per orbit of the group's operations
        mats = [np.rint(M).astype(int) for M in g["opmat"]]
        orb = lambda p: frozenset(tuple(int(t) for t in M @ np.array(p)) for M in mats)  # noqa: E731
        orbits = {orb(p) for p in lattice}
        assert len(got) == len(orbits)
        assert {orb(p) for p in got} == orbits
    else:
        assert set(got) == lattice
    # interior points: radius <= Rsort
    assert np.all(rad[:npt] <= R + EPSPOS) and np.all(rad[npt:] > R + EPSPOS)
    # boundary lists from the geometry: interior points with an outside
    # neighbour along the axes, then outside points with an interior neighbour
    nsf = int(g["nsortf"])
    ivi = g["ivois"]
    if "group" not in g:
        inner = [i for i in range(npt) if np.any(ivi[i, :6] >= npt)]
        outer = [i for i in range(npt, len(ind)) if np.any((ivi[i] >= 0) & (ivi[i] < npt))]
        assert list(g["isrt"][:nsf]) == inner and list(g["isrt"][nsf:]) == outer
    # weights: 4 pi rvol / sum(rvol) over the inner boundary points
    rvol = g["rvol"] if "rvol" in g else np.ones(npt)
    w = 4 * np.pi * rvol[g["isrt"][:nsf]] / rvol[g["isrt"][:nsf]].sum()
    assert np.allclose(g["poidso"], w, rtol=1e-14)
    assert abs(g["poidso"].sum() - 4 * np.pi) <= 1e-12
