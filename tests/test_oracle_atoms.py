"""The oracle's muffin-tin atoms against the reference (tests/golden/
atoms_R1.9_h0.25.npz, written by make_golden.py from oracle/_ref): three
atoms, two of them at the same distance from the absorber.

- structure: points inside each sphere (reseau, general.f90:9563-9577), the
  atom border lists and weights (bordure, :9924-10068), Ylmato (Ylmpt,
  :11533-11568), the atom blocks in the row order and the profile (newind,
  mat.f90:1790-1875), all equal;
- the solution of mat_solve with the reference calcMatRow (grid rows with a
  neighbour inside an atom, mat.f90:1233-1309; atom rows, :1515-1632) for a
  synthetic phiato, at every third row: bit for bit on the real branches,
  within 1e-13 with complex energy (the csin/ccos of the outer-sphere
  functions differ in their last bits)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atoms_R1.9_h0.25.npz")
BRANCHES = [(0.7, 0.05, True, True, False), (0.7, 0.0, False, False, True), (0.7, 0.0, False, True, False)]


def _cluster(g):
    return O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]),
                     atoms=(g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"]))


def phiato(c, posi_bohr, lmax, E, Eimag, E_comp):
    """j_L(k r) Y_Lm at the border points, k = sqrt(E + 1.5 (+ i Eimag))
    (the fixture's synthetic radial solutions of a flat well)."""
    from scipy.special import spherical_jn
    k = np.sqrt(E + 1.5 + (1j * Eimag if E_comp else 0))
    xyz = c.xyz
    out = []
    for ia in range(1, c.natome + 1):
        a = c.atom(ia)
        r = np.linalg.norm(xyz[a["ibord"] - 1, :3] - posi_bohr[ia - 1], axis=1)
        ls = np.array([L for L in range(int(lmax[ia - 1]) + 1) for m in range(-L, L + 1)])
        phi = np.stack([spherical_jn(L, k * r) for L in ls], axis=1) * a["ylmato"]
        out.append(phi if E_comp else phi.real)
    return out


def test_atom_structure_equals_reference():
    g = dict(np.load(G))
    c = _cluster(g)
    assert (c.npoint, c.npr, c.nligne) == (g["npoint"], g["npr"], g["nligne"])
    assert np.array_equal(c.numia, g["numia"])
    for ia in range(1, c.natome + 1):
        a = c.atom(ia)
        assert (a["nbord"], a["nbordf"], a["ianew"]) == (g["nbord"][ia - 1], g["nbordf"][ia - 1], g["ianew"][ia - 1])
        assert np.array_equal(a["ibord"], g[f"ibord{ia - 1}"])
        assert np.array_equal(a["poidsa"][: a["nbordf"]], g[f"poidsa{ia - 1}"])
        assert np.array_equal(a["ylmato"], g[f"ylmato{ia - 1}"])
        assert abs(a["poidsa"].sum() - 4 * np.pi) <= 1e-12  # bordure: 4 pi rvol / sum(rvol)
    lb1, lb2, newinv, newp = c.profile()
    assert np.array_equal(lb1, g["lb1"]) and np.array_equal(lb2, g["lb2"])
    assert np.array_equal(newinv, g["newinv"]) and np.array_equal(newp, g["new"])
    # every atom row block sits in the radial order: its first row follows the
    # last grid row closer to the absorber than the atom
    assert np.all(np.diff(lb1) >= 0)


@pytest.mark.parametrize("branch", range(3))
def test_atom_solve_equals_reference(branch):
    g = dict(np.load(G))
    E, Ei, E_comp, cal, basereel = BRANCHES[branch]
    c = _cluster(g)
    c.set_phiato(phiato(c, g["posi"] / BOHR, g["lmax"], E, Ei, E_comp))
    st, x, fr, _, _, _ = c.solve(g["V"], E, Ei, E, E_comp, cal, basereel)
    assert st == 0, fr
    ref = g[f"x{branch}"]
    got = x[g["rows"]]
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print("branch", BRANCHES[branch], "rel diff vs reference", rel)
    assert rel <= (1e-13 if E_comp else 0.0), rel
