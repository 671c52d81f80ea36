"""The rest of the reference's post-solve chain in the oracle, against the
reference's own mat routine (mat.f90:5, run from oracle/_ref by
make_golden.py into tests/golden/atoms_tau_R1.9_h0.25.npz) on the
three-atom cluster of atoms_R1.9_h0.25.npz:

- the irregular cluster solution (FDM_comp, mat.f90:135-217 with
  Test_reg = 0): newind's rows with two blocks per atom and no outer-sphere
  rows (mat.f90:1761-1921), calcMatRow's irregular branches (grid rows with
  atom neighbours :1266-1307 and outer neighbours :1318-1398, atom rows
  :1515-1630 with the border points split in halves), solved by mat_solve:
  layout equal, solution within 1e-12 (complex energy: the csin/ccos of the
  outer-sphere functions differ in their last bits);
- Tau_RI (:229-295): the bilinear contraction of the regular and irregular
  solutions' atom rows, with the regular solution rounded to single precision
  as mat.f90:211 stores it (cmplx of default kind): equal within 1e-12; the
  unrounded contraction is ~1e-8 away, so the rounding is pinned;
- Taull for real energies (:297-405) followed by mat's symmetrisation
  (:596-641, real harmonics: i Im(Taull)), real and complex basis: equal."""
import os

import numpy as np

from oracle import oracle as O

BOHR = 0.52917721067
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GA = os.path.join(HERE, "atoms_R1.9_h0.25.npz")
GT = os.path.join(HERE, "atoms_tau_R1.9_h0.25.npz")


def _load():
    return dict(np.load(GA)), dict(np.load(GT))


def _cluster(g, irregular=False):
    return O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]),
                     atoms=(g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"]), irregular=irregular)


def _split(flat, c):
    """Per-atom (nbord, nlmsa) arrays from the fixture's concatenation."""
    out, off = [], 0
    for ia in range(1, c.natome + 1):
        a = c.atom(ia)
        n = a["nbord"] * a["nlmsa"]
        out.append(flat[off:off + n].reshape(a["nbord"], a["nlmsa"]))
        off += n
    return out


def _regular_and_irregular(g, t):
    c = _cluster(g)
    ci = _cluster(g, irregular=True)
    phi, phi2 = _split(t["phi_c"], c), _split(t["phi2_c"], c)
    c.set_phiato(phi)
    ci.set_phiato(phi, phi2)
    E, Ei = float(t["Ec"]), float(t["Eimag"])
    st, xr = c.solve(t["V"], E, Ei, E, True, True, False)[:2]
    assert st == 0
    st, xi = ci.solve(t["V"], E, Ei, E, True, True, False)[:2]
    assert st == 0
    return c, ci, xr, xi


def test_irregular_layout_equals_reference():
    g, t = _load()
    ci = _cluster(g, irregular=True)
    assert ci.nligne == int(t["irr_nligne"]) == ci.npr + 2 * sum((int(L) + 1) ** 2 for L in g["lmax"])
    lb1, lb2, newinv, newp = ci.profile()
    assert np.array_equal(lb1, t["irr_lb1"]) and np.array_equal(lb2, t["irr_lb2"])
    assert np.array_equal(newinv, t["irr_newinv"]) and np.array_equal(newp, t["irr_new"])
    assert [ci.atom(ia)["ianew"] for ia in range(1, ci.natome + 1)] == list(t["irr_ianew"])
    assert not np.any(newinv == 0)  # no outer-sphere rows


def test_irregular_solution_equals_reference():
    g, t = _load()
    _, _, _, xi = _regular_and_irregular(g, t)
    ref = t["x_irr"]
    assert np.abs(xi[t["irr_rows"]] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tau_ri_equals_reference():
    g, t = _load()
    c, ci, xr, xi = _regular_and_irregular(g, t)
    ir = [c.atom(ia)["ianew"] for ia in range(1, c.natome + 1)]
    ii = [ci.atom(ia)["ianew"] for ia in range(1, c.natome + 1)]
    nlmsa = [c.atom(ia)["nlmsa"] for ia in range(1, c.natome + 1)]
    ref = t["tau_ri"]  # (natome, n_RI, nlmmax, nlmmax)
    scale = np.abs(ref).max()
    tri = O.tau_ri(xr, xi, ir, ii, nlmsa)
    for ia, na in enumerate(nlmsa):
        assert np.abs(tri[ia] - ref[ia][:, :na, :na]).max() <= 1e-12 * scale
    # the reference's single-precision copy of the regular solution is what
    # these values carry: unrounded they are 1e-9 .. 1e-7 away
    raw = O.tau_ri(xr, xi, ir, ii, nlmsa, round_regular=False)
    assert max(np.abs(raw[ia] - ref[ia][:, :na, :na]).max() for ia, na in enumerate(nlmsa)) > 1e-9 * scale
    # a contraction with the roles of the two solutions exchanged fails
    swapped = O.tau_ri(xi, xr, ii, ir, nlmsa, n_ato_ri=1)
    assert max(np.abs(swapped[ia][0] - ref[ia][0, :na, :na]).max() for ia, na in enumerate(nlmsa)) > 1e-3 * scale


def test_taull_chain_equals_reference_mat():
    g, t = _load()
    E = float(t["E"])
    for basereel in (True, False):
        c = _cluster(g)
        c.set_phiato(_split(t["phi_real"], c))
        st, x = c.solve(t["V"], E, 0.0, E, False, not basereel, basereel)[:2]
        assert st == 0
        atoms = [(c.atom(ia)["ianew"], c.atom(ia)["nlmsa"]) for ia in range(1, c.natome + 1)]
        taus, _, info = O.tau(x, c.nligneso, basereel, False, atoms)
        assert info == 0
        sym = O.tau_sym(taus, False)
        ref = t[f"taull_basereel{int(basereel)}"]
        scale = np.abs(ref).max()
        for ia, (_, na) in enumerate(atoms):
            assert np.abs(sym[ia] - ref[ia][:na, :na]).max() <= 1e-13 * scale
            assert np.abs(sym[ia].real).max() == 0.0  # real harmonics: i Im(Taull)
        # before the symmetrisation the real parts are not zero
        assert max(np.abs(tt.real).max() for tt in taus) > 1e-6 * scale


def test_tau_sym_complex_harmonics_pairs():
    """Ylm_comp branch (mat.f90:630-634): every pair and its partner
    (L2 -m2, L1 -m1) end up related by (-1)^(m1+m2), and a tau that already
    satisfies the relation is unchanged."""
    rng = np.random.default_rng(3)
    na = 9
    T = rng.normal(size=(na, na)) + 1j * rng.normal(size=(na, na))
    S = O.tau_sym([T], True)[0]
    lm = [(L, m) for L in range(3) for m in range(-L, L + 1)]
    idx = {v: k for k, v in enumerate(lm)}
    for a, (L1, m1) in enumerate(lm):
        for b, (L2, m2) in enumerate(lm):
            p, q = idx[(L2, -m2)], idx[(L1, -m1)]
            assert abs(S[a, b] - (-1) ** (m1 + m2) * S[p, q]) <= 1e-15 * abs(S).max()
    assert np.array_equal(O.tau_sym([S], True)[0], S)
